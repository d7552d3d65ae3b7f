/* Collision stages over a voxel grid: CPU restatement.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the device
 * stages behind dl_voxel_fill / dl_voxel_carve / dl_voxel_mesh /
 * dl_voxel_query, never shipped or measured.
 *
 * The reference has no code for these stages. README.md:108 and
 * src/voxel/block-mask-buffer.ts:10-11 name fill, carve and mesh as grid
 * arithmetic downstream of BlockMaskBuffer, and src/voxel/voxelize.ts:12-13
 * calls the voxelizer's output "the sparse occupancy bitmap that downstream
 * collision stages (fill, carve, mesh) operate on". The definitions below
 * are therefore this library's own:
 *
 * Grid. nb[3] blocks of 4x4x4 voxels; voxel (x, y, z) lives in block
 * (x>>2) + (y>>2)*nb0 + (z>>2)*nb0*nb1 at bit (x&3) + 4*(y&3) + 16*(z&3),
 * the bit layout of the voxelizer's masks (src/voxel/shaders/voxelize.ts:137).
 * Voxel i along axis k spans [bmin_k + i*res, bmin_k + (i+1)*res).
 *
 * Flood. R = the empty voxels joined through 6-connected empty voxels to a
 * seed. The output grid is the complement of R.
 *   fill:  the seeds are every empty voxel on the grid's boundary, so the
 *          result closes every cavity that solid voxels enclose;
 *   carve: the seeds are the voxels that hold the given world points
 *          (floor((p - bmin) / res) in f64; points outside the grid or in
 *          solid voxels seed nothing), so every voxel not reachable from
 *          the points becomes solid.
 *
 * Mesh. A face of direction d (+x, -x, +y, -y, +z, -z, in that order) is a
 * solid voxel whose neighbour along d is empty or outside the grid. For
 * normal axis a the in-plane axes (u, v) are (y, z) for x, (x, z) for y and
 * (x, y) for z. In each slice (the voxel coordinate along a) and row v the
 * faces form maximal runs [u0, u1) along u. A run is the top of a quad
 * unless row v-1 holds exactly the same run; the quad then extends over
 * every following row that holds exactly that run. Quads are ordered by
 * (d, slice, v0, u0). Each quad has 4 vertices (u0,v0), (u1,v0), (u1,v1),
 * (u0,v1) on the plane slice + (d positive ? 1 : 0), each coordinate
 * (float)(bmin_k + (double)i * res), and two triangles wound
 * counter-clockwise seen from the side d points to.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int solid_at(const uint64_t *mask, const int64_t nb[3], int64_t x, int64_t y, int64_t z) {
    if (x < 0 || y < 0 || z < 0 || x >= 4 * nb[0] || y >= 4 * nb[1] || z >= 4 * nb[2]) return 0;
    const uint64_t blk = (uint64_t)(x >> 2) + (uint64_t)(y >> 2) * (uint64_t)nb[0] +
                         (uint64_t)(z >> 2) * (uint64_t)nb[0] * (uint64_t)nb[1];
    const int bit = (int)((x & 3) + 4 * (y & 3) + 16 * (z & 3));
    return (int)((mask[blk] >> bit) & 1u);
}

/* voxel of a world point, or 0 when it lies outside the grid (or is NaN) */
static int point_voxel(const double p[3], const double bmin[3], double res, const int64_t nb[3], int64_t v[3]) {
    for (int k = 0; k < 3; k++) {
        const double f = floor((p[k] - bmin[k]) / res);
        if (!(f >= 0.0 && f < (double)(4 * nb[k]))) return 0;
        v[k] = (int64_t)f;
    }
    return 1;
}

static int flood(const uint64_t *mask, const int64_t nb[3], int outside, const int64_t *seeds, uint64_t nseeds,
                 uint64_t *out) {
    const int64_t V[3] = {4 * nb[0], 4 * nb[1], 4 * nb[2]};
    const uint64_t nvox = (uint64_t)V[0] * (uint64_t)V[1] * (uint64_t)V[2];
    const uint64_t nblk = (uint64_t)nb[0] * (uint64_t)nb[1] * (uint64_t)nb[2];
    uint8_t *reach = (uint8_t *)calloc(nvox ? nvox : 1, 1);
    uint64_t *queue = (uint64_t *)malloc((nvox ? nvox : 1) * sizeof(uint64_t));
    if (!reach || !queue) {
        free(reach);
        free(queue);
        return -1;
    }
    uint64_t head = 0, tail = 0;
#define VOX_ID(x, y, z) ((uint64_t)(x) + (uint64_t)V[0] * ((uint64_t)(y) + (uint64_t)V[1] * (uint64_t)(z)))
#define PUSH(x, y, z)                                                          \
    do {                                                                       \
        if (!solid_at(mask, nb, (x), (y), (z))) {                              \
            const uint64_t id_ = VOX_ID(x, y, z);                              \
            if (!reach[id_]) {                                                 \
                reach[id_] = 1;                                                \
                queue[tail++] = id_;                                           \
            }                                                                  \
        }                                                                      \
    } while (0)
    if (outside) {
        for (int64_t z = 0; z < V[2]; z++)
            for (int64_t y = 0; y < V[1]; y++)
                for (int64_t x = 0; x < V[0]; x++)
                    if (x == 0 || y == 0 || z == 0 || x == V[0] - 1 || y == V[1] - 1 || z == V[2] - 1)
                        PUSH(x, y, z);
    }
    for (uint64_t i = 0; i < nseeds; i++) PUSH(seeds[3 * i], seeds[3 * i + 1], seeds[3 * i + 2]);
    while (head < tail) {
        const uint64_t id = queue[head++];
        const int64_t x = (int64_t)(id % (uint64_t)V[0]);
        const int64_t y = (int64_t)((id / (uint64_t)V[0]) % (uint64_t)V[1]);
        const int64_t z = (int64_t)(id / ((uint64_t)V[0] * (uint64_t)V[1]));
        if (x > 0) PUSH(x - 1, y, z);
        if (x + 1 < V[0]) PUSH(x + 1, y, z);
        if (y > 0) PUSH(x, y - 1, z);
        if (y + 1 < V[1]) PUSH(x, y + 1, z);
        if (z > 0) PUSH(x, y, z - 1);
        if (z + 1 < V[2]) PUSH(x, y, z + 1);
    }
#undef PUSH
    for (uint64_t b = 0; b < nblk; b++) out[b] = ~(uint64_t)0;
    for (int64_t z = 0; z < V[2]; z++)
        for (int64_t y = 0; y < V[1]; y++)
            for (int64_t x = 0; x < V[0]; x++)
                if (reach[VOX_ID(x, y, z)]) {
                    const uint64_t blk = (uint64_t)(x >> 2) + (uint64_t)(y >> 2) * (uint64_t)nb[0] +
                                         (uint64_t)(z >> 2) * (uint64_t)nb[0] * (uint64_t)nb[1];
                    out[blk] &= ~((uint64_t)1 << ((x & 3) + 4 * (y & 3) + 16 * (z & 3)));
                }
#undef VOX_ID
    free(reach);
    free(queue);
    return 0;
}

int orc_vox_fill(const uint64_t *mask, const int64_t nb[3], uint64_t *out) {
    return flood(mask, nb, 1, NULL, 0, out);
}

int orc_vox_carve(const uint64_t *mask, const int64_t nb[3], const double bmin[3], double res, const double *points,
                  uint64_t npoints, uint64_t *out) {
    int64_t *seeds = (int64_t *)malloc((npoints ? npoints : 1) * 3 * sizeof(int64_t));
    if (!seeds) return -1;
    uint64_t ns = 0;
    for (uint64_t i = 0; i < npoints; i++) {
        int64_t v[3];
        if (point_voxel(points + 3 * i, bmin, res, nb, v)) {
            seeds[3 * ns] = v[0];
            seeds[3 * ns + 1] = v[1];
            seeds[3 * ns + 2] = v[2];
            ns++;
        }
    }
    const int rc = flood(mask, nb, 0, seeds, ns, out);
    free(seeds);
    return rc;
}

void orc_vox_query(const uint64_t *mask, const int64_t nb[3], const double bmin[3], double res,
                   const double *points, uint64_t npoints, uint8_t *out) {
    for (uint64_t i = 0; i < npoints; i++) {
        int64_t v[3];
        out[i] = point_voxel(points + 3 * i, bmin, res, nb, v) ? (uint8_t)solid_at(mask, nb, v[0], v[1], v[2]) : 0;
    }
}

/* axis a, in-plane axes (u, v) */
static const int kUAxis[3] = {1, 0, 0}, kVAxis[3] = {2, 2, 1};

static int face_at(const uint64_t *mask, const int64_t nb[3], int d, int64_t s, int64_t u, int64_t v) {
    const int a = d >> 1, sg = (d & 1) ? -1 : 1;
    int64_t c[3];
    c[a] = s;
    c[kUAxis[a]] = u;
    c[kVAxis[a]] = v;
    if (!solid_at(mask, nb, c[0], c[1], c[2])) return 0;
    c[a] += sg;
    return !solid_at(mask, nb, c[0], c[1], c[2]);
}

/* row v holds exactly the run [u0, u1) */
static int exact_run(const uint64_t *mask, const int64_t nb[3], int d, int64_t s, int64_t v, int64_t u0, int64_t u1,
                     int64_t Vu) {
    if (u0 > 0 && face_at(mask, nb, d, s, u0 - 1, v)) return 0;
    if (u1 < Vu && face_at(mask, nb, d, s, u1, v)) return 0;
    for (int64_t u = u0; u < u1; u++)
        if (!face_at(mask, nb, d, s, u, v)) return 0;
    return 1;
}

uint64_t orc_vox_mesh(const uint64_t *mask, const int64_t nb[3], const double bmin[3], double res, float *pos,
                      uint32_t *idx, uint64_t cap_quads) {
    const int64_t V[3] = {4 * nb[0], 4 * nb[1], 4 * nb[2]};
    uint64_t q = 0;
    for (int d = 0; d < 6; d++) {
        const int a = d >> 1, ua = kUAxis[a], va = kVAxis[a];
        const int positive = !(d & 1);
        /* u x v points along +a for x and z, along -a for y */
        const int ccw = (a == 1) ? !positive : positive;
        for (int64_t s = 0; s < V[a]; s++)
            for (int64_t v = 0; v < V[va]; v++) {
                int64_t u = 0;
                while (u < V[ua]) {
                    if (!face_at(mask, nb, d, s, u, v)) {
                        u++;
                        continue;
                    }
                    const int64_t u0 = u;
                    while (u < V[ua] && face_at(mask, nb, d, s, u, v)) u++;
                    const int64_t u1 = u;
                    if (v > 0 && exact_run(mask, nb, d, s, v - 1, u0, u1, V[ua])) continue;
                    int64_t v1 = v + 1;
                    while (v1 < V[va] && exact_run(mask, nb, d, s, v1, u0, u1, V[ua])) v1++;
                    if (q < cap_quads && pos && idx) {
                        const int64_t cu[4] = {u0, u1, u1, u0}, cv[4] = {v, v, v1, v1};
                        for (int k = 0; k < 4; k++) {
                            int64_t c[3];
                            c[a] = s + (positive ? 1 : 0);
                            c[ua] = cu[k];
                            c[va] = cv[k];
                            for (int j = 0; j < 3; j++)
                                pos[12 * q + 3 * k + j] = (float)(bmin[j] + (double)c[j] * res);
                        }
                        const uint32_t b = (uint32_t)(4 * q);
                        const uint32_t t_ccw[6] = {b, b + 1, b + 2, b, b + 2, b + 3};
                        const uint32_t t_cw[6] = {b, b + 2, b + 1, b, b + 3, b + 2};
                        memcpy(idx + 6 * q, ccw ? t_ccw : t_cw, sizeof(t_ccw));
                    }
                    q++;
                }
            }
    }
    return q;
}
