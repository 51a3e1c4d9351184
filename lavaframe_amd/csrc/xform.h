// Matrix inverses used to pre-transform instances on the host.
//
// The reference inverts the instance matrix at every TLAS-leaf visit
// (closest_hit.glsl:159-160) and the normal matrix at every hit
// (pathtrace.glsl:31).  Both are pure functions of the stored transform, so
// they are evaluated once per instance here with one fixed formula (adjugate
// over determinant, cofactors by 3x3 / 2x2 expansion along the first row,
// inverse = adj * (1 / det)); the CPU oracle evaluates the same formula per
// visit and gets the same bits.
#pragma once

namespace lfh {

// A(r, c) of a column-major m[c][r].
inline float mat_at(const float m[4][4], int r, int c) { return m[c][r]; }

inline float det3(float a00, float a01, float a02, float a10, float a11, float a12, float a20, float a21,
                  float a22) {
    return a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) + a02 * (a10 * a21 - a11 * a20);
}

// inv[c][r] (column-major) of the 4x4 m[c][r].
inline void inverse4(const float m[4][4], float inv[4][4]) {
    float C[4][4];  // cofactor C(i, j), row i, col j
    for (int i = 0; i < 4; ++i) {
        for (int j = 0; j < 4; ++j) {
            int rr[3], cc[3];
            for (int k = 0, q = 0; k < 4; ++k)
                if (k != i) rr[q++] = k;
            for (int k = 0, q = 0; k < 4; ++k)
                if (k != j) cc[q++] = k;
            float d = det3(mat_at(m, rr[0], cc[0]), mat_at(m, rr[0], cc[1]), mat_at(m, rr[0], cc[2]),
                           mat_at(m, rr[1], cc[0]), mat_at(m, rr[1], cc[1]), mat_at(m, rr[1], cc[2]),
                           mat_at(m, rr[2], cc[0]), mat_at(m, rr[2], cc[1]), mat_at(m, rr[2], cc[2]));
            C[i][j] = ((i + j) & 1) ? -d : d;
        }
    }
    float det = mat_at(m, 0, 0) * C[0][0] + mat_at(m, 0, 1) * C[0][1] + mat_at(m, 0, 2) * C[0][2] +
                mat_at(m, 0, 3) * C[0][3];
    float s = 1.0f / det;
    // inv(r, c) = C(c, r) * s, stored column-major inv[c][r].
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[c][r] = C[c][r] * s;
}

// transpose(inverse(mat3(m))) as rows: nm[r][c] = C3(r, c) / det3, where C3
// are the 2x2 cofactors of the upper-left 3x3 block.
inline void normal_matrix(const float m[4][4], float nm[3][3]) {
    float C[3][3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            int rr[2], cc[2];
            for (int k = 0, q = 0; k < 3; ++k)
                if (k != i) rr[q++] = k;
            for (int k = 0, q = 0; k < 3; ++k)
                if (k != j) cc[q++] = k;
            float d = mat_at(m, rr[0], cc[0]) * mat_at(m, rr[1], cc[1]) - mat_at(m, rr[0], cc[1]) * mat_at(m, rr[1], cc[0]);
            C[i][j] = ((i + j) & 1) ? -d : d;
        }
    }
    float det = mat_at(m, 0, 0) * C[0][0] + mat_at(m, 0, 1) * C[0][1] + mat_at(m, 0, 2) * C[0][2];
    float s = 1.0f / det;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) nm[r][c] = C[r][c] * s;
}

}  // namespace lfh
