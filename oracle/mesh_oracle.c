/*
 * CPU oracle: surface nets on the brick TSDF and the mesh post-processing of the reference.
 * TEST INFRASTRUCTURE ONLY (see lr3d_oracle.h).
 *
 * Restates frontend/src/app/vision/fusion/surface-nets.ts (SN:<line>) and
 * frontend/src/app/vision/fusion/mesh.ts (MS:<line>).  JavaScript numbers are doubles: positions
 * and edge crossings are formed in double and stored as float (Float32Array.from); the diagonal
 * choice compares the double positions; normals accumulate face by face into a Float32Array
 * (each += rounds to float); vector length is sqrt(x^2 + y^2 + z^2) in double (the reference's
 * Math.hypot is left to the JavaScript engine, so its last-ulp behaviour is parity unpinned).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lr3d_oracle.h"

/* SN:17-41 */
static const int CORNERS[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
static const int EDGES[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                                 {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

typedef struct {
    const orc_sdf_bricks *v;
    int32_t extent[3];
} sn_ctx;

/* SN:53-58 */
static int64_t slot(const sn_ctx *c, int32_t x, int32_t y, int32_t z) {
    if (x < 0 || y < 0 || z < 0 || x >= c->extent[0] || y >= c->extent[1] || z >= c->extent[2]) return -1;
    const int32_t *d = c->v->dims;
    int32_t brick = c->v->lookup[((z >> 3) * d[1] + (y >> 3)) * d[0] + (x >> 3)];
    if (brick < 0) return -1;
    return (int64_t)brick * 512 + (x & 7) + ((y & 7) << 3) + ((z & 7) << 6);
}

/* SN:76-84 */
static int sample_cell(const sn_ctx *c, double min_weight, int32_t x, int32_t y, int32_t z, float values[8]) {
    int inside = 0;
    for (int k = 0; k < 8; k++) {
        int64_t s = slot(c, x + CORNERS[k][0], y + CORNERS[k][1], z + CORNERS[k][2]);
        if (s < 0 || (double)c->v->weight[s] < min_weight) return 0;
        values[k] = c->v->tsdf[s];
        if (values[k] < 0) inside++;
    }
    return inside > 0 && inside < 8;
}

static double dist2(const double *p, uint32_t a, uint32_t b) {
    double dx = p[a * 3] - p[b * 3], dy = p[a * 3 + 1] - p[b * 3 + 1], dz = p[a * 3 + 2] - p[b * 3 + 2];
    return dx * dx + dy * dy + dz * dz;
}

/* MS:9-32 computeNormals (float accumulation, double arithmetic per step) */
void orc_compute_normals(const float *p, const uint32_t *indices, int64_t n_indices, int64_t n_vertices,
                         float *normals) {
    memset(normals, 0, sizeof(float) * 3 * (size_t)n_vertices);
    for (int64_t f = 0; f < n_indices; f += 3) {
        size_t a = (size_t)indices[f] * 3, b = (size_t)indices[f + 1] * 3, c = (size_t)indices[f + 2] * 3;
        double e1[3] = {(double)p[b] - p[a], (double)p[b + 1] - p[a + 1], (double)p[b + 2] - p[a + 2]};
        double e2[3] = {(double)p[c] - p[a], (double)p[c + 1] - p[a + 1], (double)p[c + 2] - p[a + 2]};
        double nx = e1[1] * e2[2] - e1[2] * e2[1];
        double ny = e1[2] * e2[0] - e1[0] * e2[2];
        double nz = e1[0] * e2[1] - e1[1] * e2[0];
        size_t vs[3] = {a, b, c};
        for (int k = 0; k < 3; k++) {
            normals[vs[k]] = (float)((double)normals[vs[k]] + nx);
            normals[vs[k] + 1] = (float)((double)normals[vs[k] + 1] + ny);
            normals[vs[k] + 2] = (float)((double)normals[vs[k] + 2] + nz);
        }
    }
    for (int64_t v = 0; v < n_vertices * 3; v += 3) {
        double x = normals[v], y = normals[v + 1], z = normals[v + 2];
        double length = sqrt(x * x + y * y + z * z);
        if (length == 0.0) length = 1.0;
        normals[v] = (float)(x / length);
        normals[v + 1] = (float)(y / length);
        normals[v + 2] = (float)(z / length);
    }
}

/* SN:47-166 extractSurface.  Outputs are malloc'ed (free with orc_mesh_free). */
int32_t orc_extract_surface(const orc_sdf_bricks *vol, double min_weight, orc_mesh *out) {
    memset(out, 0, sizeof(*out));
    sn_ctx c = {vol, {vol->dims[0] * 8, vol->dims[1] * 8, vol->dims[2] * 8}};
    const int64_t cells = (int64_t)vol->n_bricks * 512;
    int64_t *cell_vertex = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cells > 0 ? cells : 1));
    int32_t *brick_coord = (int32_t *)calloc((size_t)vol->n_bricks * 3 + 1, sizeof(int32_t));
    size_t cap = 1024, n_pos = 0;
    double *pos = (double *)malloc(sizeof(double) * cap * 3);
    if (!cell_vertex || !brick_coord || !pos) return -1;
    for (int64_t i = 0; i < cells; i++) cell_vertex[i] = -1;
    /* SN:67-74 */
    const int64_t grid = (int64_t)vol->dims[0] * vol->dims[1] * vol->dims[2];
    for (int64_t i = 0; i < grid; i++) {
        int32_t b = vol->lookup[i];
        if (b < 0) continue;
        brick_coord[b * 3] = (int32_t)(i % vol->dims[0]);
        brick_coord[b * 3 + 1] = (int32_t)((i / vol->dims[0]) % vol->dims[1]);
        brick_coord[b * 3 + 2] = (int32_t)(i / ((int64_t)vol->dims[0] * vol->dims[1]));
    }
    float values[8];
    /* SN:87-115 pass 1: vertices */
    for (int32_t b = 0; b < vol->n_bricks; b++) {
        for (int32_t local = 0; local < 512; local++) {
            int32_t x = brick_coord[b * 3] * 8 + (local & 7);
            int32_t y = brick_coord[b * 3 + 1] * 8 + ((local >> 3) & 7);
            int32_t z = brick_coord[b * 3 + 2] * 8 + (local >> 6);
            if (!sample_cell(&c, min_weight, x, y, z, values)) continue;
            double sx = 0, sy = 0, sz = 0;
            int n = 0;
            for (int e = 0; e < 12; e++) {
                int a = EDGES[e][0], q = EDGES[e][1];
                double va = values[a], vc = values[q];
                if ((va < 0) == (vc < 0)) continue;
                double t = va / (va - vc);
                sx += CORNERS[a][0] + (CORNERS[q][0] - CORNERS[a][0]) * t;
                sy += CORNERS[a][1] + (CORNERS[q][1] - CORNERS[a][1]) * t;
                sz += CORNERS[a][2] + (CORNERS[q][2] - CORNERS[a][2]) * t;
                n++;
            }
            if (n_pos == cap) {
                cap *= 2;
                pos = (double *)realloc(pos, sizeof(double) * cap * 3);
                if (!pos) return -1;
            }
            cell_vertex[(int64_t)b * 512 + local] = (int64_t)n_pos;
            pos[n_pos * 3] = vol->origin[0] + (x + 0.5 + sx / n) * vol->voxel;
            pos[n_pos * 3 + 1] = vol->origin[1] + (y + 0.5 + sy / n) * vol->voxel;
            pos[n_pos * 3 + 2] = vol->origin[2] + (z + 0.5 + sz / n) * vol->voxel;
            n_pos++;
        }
    }
    /* SN:117-156 pass 2: quads */
    static const int AXES[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    size_t icap = 1024, n_idx = 0;
    uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * icap);
    if (!idx) return -1;
    for (int32_t b = 0; b < vol->n_bricks; b++) {
        for (int32_t local = 0; local < 512; local++) {
            int32_t x = brick_coord[b * 3] * 8 + (local & 7);
            int32_t y = brick_coord[b * 3 + 1] * 8 + ((local >> 3) & 7);
            int32_t z = brick_coord[b * 3 + 2] * 8 + (local >> 6);
            int64_t s0 = (int64_t)b * 512 + local;
            if ((double)vol->weight[s0] < min_weight) continue;
            float v0 = vol->tsdf[s0];
            for (int a = 0; a < 3; a++) {
                int64_t s1 = slot(&c, x + AXES[a][0], y + AXES[a][1], z + AXES[a][2]);
                if (s1 < 0 || (double)vol->weight[s1] < min_weight) continue;
                float v1 = vol->tsdf[s1];
                if ((v0 < 0) == (v1 < 0)) continue;
                const int *u = AXES[(a + 1) % 3], *w = AXES[(a + 2) % 3];
                int64_t qs[4];
                int64_t sq[4] = {slot(&c, x - u[0] - w[0], y - u[1] - w[1], z - u[2] - w[2]),
                                 slot(&c, x - w[0], y - w[1], z - w[2]), slot(&c, x, y, z),
                                 slot(&c, x - u[0], y - u[1], z - u[2])};
                int ok = 1;
                for (int k = 0; k < 4; k++) {
                    qs[k] = sq[k] < 0 ? -1 : cell_vertex[sq[k]];
                    if (qs[k] < 0) ok = 0;
                }
                if (!ok) continue;
                uint32_t quad[4];
                if (v0 < 0) {
                    quad[0] = (uint32_t)qs[0]; quad[1] = (uint32_t)qs[1]; quad[2] = (uint32_t)qs[2]; quad[3] = (uint32_t)qs[3];
                } else {
                    quad[0] = (uint32_t)qs[0]; quad[1] = (uint32_t)qs[3]; quad[2] = (uint32_t)qs[2]; quad[3] = (uint32_t)qs[1];
                }
                if (n_idx + 6 > icap) {
                    icap *= 2;
                    idx = (uint32_t *)realloc(idx, sizeof(uint32_t) * icap);
                    if (!idx) return -1;
                }
                if (dist2(pos, quad[0], quad[2]) <= dist2(pos, quad[1], quad[3])) {
                    uint32_t t[6] = {quad[0], quad[1], quad[2], quad[0], quad[2], quad[3]};
                    memcpy(idx + n_idx, t, sizeof(t));
                } else {
                    uint32_t t[6] = {quad[0], quad[1], quad[3], quad[1], quad[2], quad[3]};
                    memcpy(idx + n_idx, t, sizeof(t));
                }
                n_idx += 6;
            }
        }
    }
    out->n_vertices = (int64_t)n_pos;
    out->n_indices = (int64_t)n_idx;
    out->positions = (float *)malloc(sizeof(float) * 3 * (n_pos ? n_pos : 1));
    out->normals = (float *)malloc(sizeof(float) * 3 * (n_pos ? n_pos : 1));
    out->indices = idx;
    for (size_t i = 0; i < n_pos * 3; i++) out->positions[i] = (float)pos[i];
    orc_compute_normals(out->positions, idx, (int64_t)n_idx, (int64_t)n_pos, out->normals);
    free(pos);
    free(cell_vertex);
    free(brick_coord);
    return 0;
}

static int64_t find_root(int64_t *parent, int64_t i) {
    while (parent[i] != i) {
        parent[i] = parent[parent[i]];
        i = parent[i];
    }
    return i;
}

/* MS:38-90 removeSmallComponents */
int32_t orc_remove_small_components(const orc_mesh *in, double fraction, int32_t min_faces, orc_mesh *out) {
    memset(out, 0, sizeof(*out));
    const int64_t V = in->n_vertices, N = in->n_indices;
    int64_t *parent = (int64_t *)malloc(sizeof(int64_t) * (size_t)(V ? V : 1));
    int64_t *faces = (int64_t *)calloc((size_t)(V ? V : 1), sizeof(int64_t));
    int64_t *remap = (int64_t *)malloc(sizeof(int64_t) * (size_t)(V ? V : 1));
    if (!parent || !faces || !remap) return -1;
    for (int64_t i = 0; i < V; i++) parent[i] = i;
    for (int64_t f = 0; f < N; f += 3) {
        int64_t a = find_root(parent, in->indices[f]);
        int64_t b = find_root(parent, in->indices[f + 1]);
        int64_t c = find_root(parent, in->indices[f + 2]);
        parent[b] = a;
        parent[find_root(parent, c)] = a;
    }
    for (int64_t f = 0; f < N; f += 3) faces[find_root(parent, in->indices[f])]++;
    const double limit = fmax((double)min_faces, (double)(N / 3) * fraction);
    for (int64_t i = 0; i < V; i++) remap[i] = -1;
    out->positions = (float *)malloc(sizeof(float) * 3 * (size_t)(V ? V : 1));
    out->normals = (float *)malloc(sizeof(float) * 3 * (size_t)(V ? V : 1));
    out->indices = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(N ? N : 1));
    if (!out->positions || !out->normals || !out->indices) return -1;
    int64_t nv = 0, ni = 0;
    for (int64_t f = 0; f < N; f += 3) {
        if ((double)faces[find_root(parent, in->indices[f])] < limit) continue;
        for (int k = 0; k < 3; k++) {
            uint32_t v = in->indices[f + k];
            if (remap[v] < 0) {
                remap[v] = nv;
                memcpy(out->positions + nv * 3, in->positions + (size_t)v * 3, 12);
                memcpy(out->normals + nv * 3, in->normals + (size_t)v * 3, 12);
                nv++;
            }
            out->indices[ni++] = (uint32_t)remap[v];
        }
    }
    out->n_vertices = nv;
    out->n_indices = ni;
    free(parent);
    free(faces);
    free(remap);
    return 0;
}

void orc_mesh_free(orc_mesh *m) {
    free(m->positions);
    free(m->normals);
    free(m->indices);
    memset(m, 0, sizeof(*m));
}
