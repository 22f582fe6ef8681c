/*
 * refcpu.c - CPU restatement ("oracle") of the sdf2mesh meshing path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library, and only as the checker or the CPU baseline;
 * the product library (libsdf2mesh.so) never links, loads or calls it.
 *
 * It restates, in plain C with fp32 arithmetic in the reference's operation order
 * (built with -ffp-contract=off, SSE math, no fast-math):
 *   ora_eval_f32 / ora_eval_f64   SDF library + tape     sdf3d_primitives.wgsl:7-36,
 *                                 sdf_op.wgsl:7-23, examples/mandelmesh.frag:3-28
 *   ora_cell_bounds               cell_bounds            dualcontour.wgsl:22-27
 *   ora_cell_vertex               cell_fetch_interpolated_pos, _cell_adapt,
 *                                 _cell_change           dualcontour.wgsl:72-131
 *   ora_cell_signs                cell_sign_changes(_f32) dualcontour.wgsl:57-69
 *   ora_normal                    sdf3d_normal + normalize sdf3d_normal.wgsl:4-9,
 *                                 dualcontour.wgsl:170-171
 *   ora_mesh                      kernel main (dualcontour.wgsl:161-180) driven by the
 *                                 per-slice loop + host scan (main.rs:294-357),
 *                                 including the one-slice readback lag
 *   ora_triangles                 VertexList::fetch_triangle_indices + vertex_index
 *                                 (mesh.rs:267-331), Quad::swap/make_triangles/is_valid
 *                                 (lib.rs:188-211)
 *   ora_write_stl / ora_write_ply STLWriter / PLYWriter (mesh.rs:8-141,155-210)
 *   ora_format_f32                Rust `{}` Display of f32 (shortest round-trip digits,
 *                                 fixed notation)
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   primitives / ops      closed-form distances at analytic points; fp32 vs fp64
 *                         agreement; smooth ops reduce to min/max far apart and to
 *                         d -/+ k/4 for equal inputs
 *   mandelbulb            closed forms: immediate bailout gives 0.5*log|p|*|p| - off;
 *                         one iteration gives 0.5*log(r^8)*r^8/(8 r^7 + 1) - off
 *   ora_cell_bounds       h*(res-1) spans the box; values equal min + h*i in fp32
 *   ora_cell_vertex       brute force over all 3^8 corner sign classes against an
 *                         independent geometric formulation (crossing points of the 12
 *                         cube edges, counted iff not at the origin); linear fields
 *                         give vertices on the plane inside the cell
 *   ora_triangles         watertight, consistently oriented meshes with Euler
 *                         characteristic 2 (sphere) and 0 (torus); 2 triangles/quad;
 *                         vertex order = ascending x|y<<16|z<<32
 *   ora_format_f32        Rust formatting known answers ("1", "0.1", "0.0000001",
 *                         "-0", "340282350000000000000000000000000000000", ...)
 *   ora_normal            sphere: normal = p/|p| within O(eps^2)
 * Parity unpinned (no reference run is possible; see the measurement notes):
 *   the fp32 bits lavapipe produces for transcendental functions (Mandelbulb), its
 *   lowering of length/normalize/mix (restated here from the WGSL/GLSL definitions),
 *   and the wgpu buffer-mapping order behind the one-slice lag (restated from
 *   main.rs:298-356 and wgpu's documented map_async semantics).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/sdf2mesh.h"

#define ORA_API __attribute__((visibility("default")))

/* Oracle-only primitive codes: the reference's example programs (examples.inc), the
 * checker for programs the front end compiles to native code. */
#define ORA_OP_MARTIN_CUBE 100u
#define ORA_OP_P_KEY 101u
#define ORA_OP_ENGRAVED_CUBE 102u

/* ---- SDF evaluators ------------------------------------------------------------- */
#define T float
#define SFX(n) n##_f32
#define SQRT sqrtf
#define FABS fabsf
#define FMIN fminf
#define FMAX fmaxf
#define ATAN atanf
#define ASIN asinf
#define POW powf
#define COS cosf
#define SIN sinf
#define LOG logf
#include "sdf_eval.inc"
#undef T
#undef SFX
#undef SQRT
#undef FABS
#undef FMIN
#undef FMAX
#undef ATAN
#undef ASIN
#undef POW
#undef COS
#undef SIN
#undef LOG

#define T double
#define SFX(n) n##_f64
#define SQRT sqrt
#define FABS fabs
#define FMIN fmin
#define FMAX fmax
#define ATAN atan
#define ASIN asin
#define POW pow
#define COS cos
#define SIN sin
#define LOG log
#include "sdf_eval.inc"

ORA_API int ora_tape_check(const s2m_op* ops, uint32_t n)
{
    int sp = 0;
    if (n == 0 || n > S2M_MAX_TAPE) return 0;
    for (uint32_t i = 0; i < n; ++i) {
        switch (ops[i].code) {
        case S2M_OP_SPHERE: case S2M_OP_BOX: case S2M_OP_TORUS: case S2M_OP_CYLINDER:
        case S2M_OP_CAPSULE: case S2M_OP_MANDELBULB: case ORA_OP_MARTIN_CUBE: case ORA_OP_P_KEY:
        case ORA_OP_ENGRAVED_CUBE:
            if (++sp > (int)S2M_MAX_STACK) return 0;
            break;
        case S2M_OP_UNION: case S2M_OP_INTERSECTION: case S2M_OP_SUBTRACTION:
        case S2M_OP_SMOOTH_UNION: case S2M_OP_SMOOTH_INTERSECTION:
        case S2M_OP_SMOOTH_SUBTRACTION:
            if (sp < 2) return 0;
            --sp;
            break;
        default: return 0;
        }
    }
    return sp == 1;
}

ORA_API void ora_eval_f32(const s2m_op* ops, uint32_t n, const float* pts, uint64_t npts,
                          float* out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)npts; ++i)
        out[i] = tape_eval_f32(ops, n, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
}

ORA_API void ora_eval_f64(const s2m_op* ops, uint32_t n, const float* pts, uint64_t npts,
                          double* out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)npts; ++i)
        out[i] = tape_eval_f64(ops, n, (double)pts[3 * i], (double)pts[3 * i + 1],
                               (double)pts[3 * i + 2]);
}

/* The fp64 Mandelbulb of mandelbulb_f64 (examples/mandelmesh.frag:3-28) with one
 * intermediate value multiplied by (1 + u): site 8k + j perturbs, in escape step k,
 * j = 0..2 the components of z entering the step, 3 its r = length(z), 4 theta * power,
 * 5 phi * power, 6 the new dr, 7 the new r = pow(r, power); sites 8 * iters + 0..2 the
 * final 0.5 * log(r), its product with r and the quotient by dr.  site < 0 perturbs
 * nothing and records the margins: *bail the smallest |r - bailout| / bailout of any
 * bailout test, *axis the smallest rho / r (rho = |z.xy|) of any escape step. */
static double mandel_site(const s2m_op* m, double px, double py, double pz, int site, double u,
                          double* bail, double* axis)
{
    const double power = m->f[0], bailout = m->f[1], offset = m->f[2];
    const uint32_t iters = m->iarg & 0xffffu;
    if (m->iarg & S2M_MANDEL_SWAP_YZ) {
        const double t = py;
        py = pz;
        pz = t;
    }
#define PERTURB(v, s) ((site == (int)(s)) ? (v) * (1.0 + u) : (v))
    double zx = px, zy = py, zz = pz, r = 0.0, dr = 1.0;
    for (uint32_t k = 0; k < iters; ++k) {
        zx = PERTURB(zx, 8 * k + 0);
        zy = PERTURB(zy, 8 * k + 1);
        zz = PERTURB(zz, 8 * k + 2);
        r = PERTURB(sqrt((zx * zx + zy * zy) + zz * zz), 8 * k + 3);
        if (site < 0) *bail = fmin(*bail, fabs(r - bailout) / bailout);
        if (r > bailout) continue;
        if (site < 0) *axis = fmin(*axis, r > 0.0 ? sqrt(zx * zx + zy * zy) / r : 0.0);
        const double theta = PERTURB(atan(zy / zx) * power, 8 * k + 4);
        const double phi = PERTURB(asin(zz / r) * power, 8 * k + 5);
        dr = PERTURB((pow(r, power - 1.0) * dr) * power + 1.0, 8 * k + 6);
        r = PERTURB(pow(r, power), 8 * k + 7);
        zx = r * (cos(theta) * cos(phi)) + px;
        zy = r * (sin(theta) * cos(phi)) + py;
        zz = r * sin(phi) + pz;
    }
    const double l = PERTURB(0.5 * log(r), 8 * iters + 0);
    const double lr = PERTURB(l * r, 8 * iters + 1);
    return PERTURB(lr / dr, 8 * iters + 2) - offset;
#undef PERTURB
}

/* Conditioning of the first Mandelbulb op of the tape at each point, three values:
 *   out[3i]     min over the executed bailout tests of |r - bailout| / bailout: a rounding
 *               of r by about this relative amount takes the other branch of
 *               `if (r > bailout) continue;` and changes the result discontinuously;
 *   out[3i + 1] min over the escape steps of rho / r: the derivatives of atan(zy / zx)
 *               and asin(zz / r) grow as r / rho (0 at r = 0);
 *   out[3i + 2] first-order error scale of an fp32 evaluation: the sum over the sites of
 *               mandel_site of |f(site perturbed by u = 2^-24) - f|, i.e. how much one
 *               relative rounding of each intermediate moves the fp64 result.
 * A tape without a Mandelbulb op gives +inf, +inf, 0. */
ORA_API void ora_mandel_cond(const s2m_op* ops, uint32_t n, const float* pts, uint64_t npts,
                             double* out)
{
    const s2m_op* m = NULL;
    for (uint32_t k = 0; k < n && !m; ++k)
        if (ops[k].code == S2M_OP_MANDELBULB) m = &ops[k];
    const double u = ldexp(1.0, -24);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)npts; ++i) {
        double bail = INFINITY, axis = INFINITY, scale = 0.0;
        if (m) {
            const double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
            const double f0 = mandel_site(m, px, py, pz, -1, u, &bail, &axis);
            const int nsite = 8 * (int)(m->iarg & 0xffffu) + 3;
            for (int s = 0; s < nsite; ++s)
                scale += fabs(mandel_site(m, px, py, pz, s, u, &bail, &axis) - f0);
        }
        out[3 * i] = bail;
        out[3 * i + 1] = axis;
        out[3 * i + 2] = scale;
    }
}

/* ---- cell geometry -------------------------------------------------------------- */
/* dualcontour.wgsl:22-27 for one axis: size = (max-min)/f32(res-1),
 * cmin = min + size*f32(i), cmax = cmin + size.  With lattice != 0 the max corner is
 * instead the next cell's min corner, min + size*f32(i+1) (shared-lattice mode). */
ORA_API void ora_cell_bounds(float bmin, float bmax, uint32_t res, uint32_t i, int lattice,
                             float* lo, float* hi)
{
    float v = (float)(uint32_t)(res - 1u);
    float size = (bmax - bmin) / v;
    float cmin = bmin + size * (float)i;
    *lo = cmin;
    *hi = lattice ? bmin + size * (float)(i + 1u) : cmin + size;
}

/* dualcontour.wgsl:57-69: bits [c100>0], [c010>0], [c001>0], [c000>0] */
ORA_API uint32_t ora_cell_signs(const float c[8])
{
    return (uint32_t)(c[1] > 0.0f) | ((uint32_t)(c[2] > 0.0f) << 1) |
           ((uint32_t)(c[4] > 0.0f) << 2) | ((uint32_t)(c[0] > 0.0f) << 3);
}

/* A NaN result is written as the canonical quiet NaN (0x7fc00000), as the library writes
 * it: the sign and payload of a NaN depend on which operand each operation propagates. */
static float canon_nan(float v) { return v != v ? NAN : v; }

/* dualcontour.wgsl:72-131.  c[] in corner order 000,100,010,110,001,101,011,111.
 * Returns 1 and writes pos when at least two crossings are counted. */
ORA_API int ora_cell_vertex(const float c[8], const float cmin[3], const float cmax[3],
                            float pos[3])
{
    const float c000 = c[0], c100 = c[1], c010 = c[2], c110 = c[3];
    const float c001 = c[4], c101 = c[5], c011 = c[6], c111 = c[7];
    float ch[12][3];
#define ADAPT(v0, v1) ((0.0f - (v0)) / ((v1) - (v0)))
#define CHANGE(i, a, b, X, Y, Z)                                                          \
    do {                                                                                  \
        float X_ = (X), Y_ = (Y), Z_ = (Z);                                               \
        if (((a) > 0.0f) != ((b) > 0.0f)) {                                               \
            ch[i][0] = X_; ch[i][1] = Y_; ch[i][2] = Z_;                                  \
        } else {                                                                          \
            ch[i][0] = 0.0f; ch[i][1] = 0.0f; ch[i][2] = 0.0f;                            \
        }                                                                                 \
    } while (0)
    /* z direction, :98-102 */
    CHANGE(0, c000, c001, 0.0f, 0.0f, ADAPT(c000, c001));
    CHANGE(1, c010, c011, 0.0f, 1.0f, ADAPT(c010, c011));
    CHANGE(2, c100, c101, 1.0f, 0.0f, ADAPT(c100, c101));
    CHANGE(3, c110, c111, 1.0f, 1.0f, ADAPT(c110, c111));
    /* y direction, :104-108 */
    CHANGE(4, c000, c010, 0.0f, ADAPT(c000, c010), 0.0f);
    CHANGE(5, c001, c011, 0.0f, ADAPT(c001, c011), 1.0f);
    CHANGE(6, c100, c110, 1.0f, ADAPT(c100, c110), 0.0f);
    CHANGE(7, c101, c111, 1.0f, ADAPT(c101, c111), 1.0f);
    /* x direction, :110-114 */
    CHANGE(8, c000, c100, ADAPT(c000, c100), 0.0f, 0.0f);
    CHANGE(9, c001, c101, ADAPT(c001, c101), 0.0f, 1.0f);
    CHANGE(10, c010, c110, ADAPT(c010, c110), 1.0f, 0.0f);
    CHANGE(11, c011, c111, ADAPT(c011, c111), 1.0f, 1.0f);
#undef CHANGE
#undef ADAPT
    /* :116-124 */
    float ax = 0.0f, ay = 0.0f, az = 0.0f, cnt = 0.0f;
    for (int i = 0; i < 12; ++i) {
        if (ch[i][0] > 0.0f || ch[i][1] > 0.0f || ch[i][2] > 0.0f) {
            ax = ax + ch[i][0];
            ay = ay + ch[i][1];
            az = az + ch[i][2];
            cnt = cnt + 1.0f;
        }
    }
    /* :126-130 */
    if (cnt <= 1.0f) return 0;
    float sx = cmax[0] - cmin[0], sy = cmax[1] - cmin[1], sz = cmax[2] - cmin[2];
    pos[0] = canon_nan(cmin[0] + (sx * ax) / cnt);
    pos[1] = canon_nan(cmin[1] + (sy * ay) / cnt);
    pos[2] = canon_nan(cmin[2] + (sz * az) / cnt);
    return 1;
}

/* sdf3d_normal.wgsl:4-9 then normalize (dualcontour.wgsl:171):
 * v1*sdf(p+v1*eps) + v2*sdf(p+v2*eps) + v3*sdf(p+v3*eps) + v4*sdf(p+v4*eps) with
 * v1=(1,-1,-1), v2=(-1,-1,1), v3=(-1,1,-1), v4=(1,1,1); normalize(v) = v/length(v). */
ORA_API void ora_normal(const s2m_op* ops, uint32_t n, const float p[3], float eps,
                        float nrm[3])
{
    static const float V[4][3] = {{1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1}};
    float d[4];
    for (int k = 0; k < 4; ++k)
        d[k] = tape_eval_f32(ops, n, p[0] + V[k][0] * eps, p[1] + V[k][1] * eps,
                             p[2] + V[k][2] * eps);
    float g[3];
    for (int a = 0; a < 3; ++a)
        g[a] = ((V[0][a] * d[0] + V[1][a] * d[1]) + V[2][a] * d[2]) + V[3][a] * d[3];
    float len = sqrtf((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
    nrm[0] = canon_nan(g[0] / len);
    nrm[1] = canon_nan(g[1] / len);
    nrm[2] = canon_nan(g[2] / len);
}

/* ---- results -------------------------------------------------------------------- */
typedef struct ora_mesh_t {
    uint64_t nvtx;
    float* vtx;       /* 6 per vertex */
    uint64_t* keys;   /* real cell coordinates x | y<<16 | z<<32 */
    uint8_t* signs;
    uint64_t ntri;
    uint32_t* tri;    /* 3 per triangle */
    uint64_t invalid_quads;
} ora_mesh_t;

ORA_API void ora_mesh_free(ora_mesh_t* m)
{
    if (!m) return;
    free(m->vtx);
    free(m->keys);
    free(m->signs);
    free(m->tri);
    free(m);
}

static uint64_t key_of(uint64_t x, uint64_t y, uint64_t z) { return x | (y << 16) | (z << 32); }

/* VertexList::vertex_index (mesh.rs:327-331): binary search by key, u32::MAX if absent */
static uint32_t vertex_index(const uint64_t* keys, uint64_t n, uint64_t key)
{
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < key) lo = mid + 1;
        else if (keys[mid] > key) hi = mid;
        else return (uint32_t)mid;
    }
    return UINT32_MAX;
}

/* mesh.rs:267-324 with Quad::swap / make_triangles / is_valid (lib.rs:188-211).
 * keys are the labels the scan gave the items (in the reference, z + 1 under the
 * readback lag), signs the 4 decoded bits.  Writes up to 6*n triangles into tri. */
static uint64_t triangles_from_list(const uint64_t* keys, const uint8_t* signs, uint64_t n,
                                    uint32_t* tri, uint64_t* invalid)
{
    uint64_t nt = 0, inv = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t x = keys[i] & 0xffff, y = (keys[i] >> 16) & 0xffff, z = (keys[i] >> 32) & 0xffff;
        int c0 = (signs[i] >> 0) & 1, c1 = (signs[i] >> 1) & 1, c2 = (signs[i] >> 2) & 1,
            c3 = (signs[i] >> 3) & 1;
        for (int e = 0; e < 3; ++e) {
            uint32_t q[4];
            int swap;
            if (e == 0) {
                if (!(c0 != c3 && y > 0 && z > 0)) continue;
                q[0] = vertex_index(keys, n, key_of(x, y - 1, z - 1));
                q[1] = vertex_index(keys, n, key_of(x, y, z - 1));
                q[2] = vertex_index(keys, n, key_of(x, y, z));
                q[3] = vertex_index(keys, n, key_of(x, y - 1, z));
                swap = c0;
            } else if (e == 1) {
                if (!(c1 != c3 && x > 0 && z > 0)) continue;
                q[0] = vertex_index(keys, n, key_of(x - 1, y, z - 1));
                q[1] = vertex_index(keys, n, key_of(x, y, z - 1));
                q[2] = vertex_index(keys, n, key_of(x, y, z));
                q[3] = vertex_index(keys, n, key_of(x - 1, y, z));
                swap = !c1;
            } else {
                if (!(c2 != c3 && x > 0 && y > 0)) continue;
                q[0] = vertex_index(keys, n, key_of(x - 1, y - 1, z));
                q[1] = vertex_index(keys, n, key_of(x, y - 1, z));
                q[2] = vertex_index(keys, n, key_of(x, y, z));
                q[3] = vertex_index(keys, n, key_of(x - 1, y, z));
                swap = c2;
            }
            if (swap) { /* Quad::swap(true) = (q3, q2, q1, q0) */
                uint32_t t0 = q[0], t1 = q[1];
                q[0] = q[3]; q[1] = q[2]; q[2] = t1; q[3] = t0;
            }
            if (q[0] == UINT32_MAX || q[1] == UINT32_MAX || q[2] == UINT32_MAX ||
                q[3] == UINT32_MAX) {
                ++inv;
                continue;
            }
            /* make_triangles: (q2, q1, q0), (q0, q3, q2) */
            uint32_t* t = tri + 3 * nt;
            t[0] = q[2]; t[1] = q[1]; t[2] = q[0];
            t[3] = q[0]; t[4] = q[3]; t[5] = q[2];
            nt += 2;
        }
    }
    *invalid = inv;
    return nt;
}

/* One z-slice of the kernel (dualcontour.wgsl:161-180): per cell of the slice, the
 * two rgba32f texels (position, flag) and (normal, signs).  tex holds 8 floats per
 * cell in (y, x) order; corners == NULL means evaluate the tape at the corners. */
static void run_slice(const s2m_op* ops, uint32_t nops, const float* corners,
                      const float* bmin, const float* bmax, const uint32_t* res, uint32_t z,
                      int lattice, float eps, int want_normals, float* tex)
{
    const int64_t nxy = (int64_t)res[0] * res[1];
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < nxy; ++c) {
        uint32_t x = (uint32_t)(c % res[0]), y = (uint32_t)(c / res[0]);
        float lo[3], hi[3], v[8], p[3];
        ora_cell_bounds(bmin[0], bmax[0], res[0], x, lattice, &lo[0], &hi[0]);
        ora_cell_bounds(bmin[1], bmax[1], res[1], y, lattice, &lo[1], &hi[1]);
        ora_cell_bounds(bmin[2], bmax[2], res[2], z, lattice, &lo[2], &hi[2]);
        if (corners) {
            memcpy(v, corners + 8 * ((uint64_t)z * nxy + c), sizeof v);
        } else { /* cell_new, dualcontour.wgsl:29-43 */
            for (int k = 0; k < 8; ++k)
                v[k] = tape_eval_f32(ops, nops, (k & 1) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1],
                                     (k & 4) ? hi[2] : lo[2]);
        }
        float* t = tex + 8 * c;
        if (ora_cell_vertex(v, lo, hi, p)) {
            float nrm[3] = {0.0f, 0.0f, 0.0f};
            if (want_normals && ops) ora_normal(ops, nops, p, eps, nrm);
            t[0] = p[0]; t[1] = p[1]; t[2] = p[2]; t[3] = 1.0f;
            t[4] = nrm[0]; t[5] = nrm[1]; t[6] = nrm[2]; t[7] = (float)ora_cell_signs(v);
        } else {
            memset(t, 0, 8 * sizeof(float));
        }
    }
}

/* The reference driver, main.rs:294-357 then TriangleMesh::from (mesh.rs:334-341).
 * With lag != 0 the host scan of iteration z reads the textures of slice z-1 and labels
 * the cells z (staging buffers are mapped before the submit of iteration z); slice
 * res-1 is computed but never read.  Keys are returned as real cell coordinates. */
static ora_mesh_t* mesh_driver(const s2m_op* ops, uint32_t nops, const float* corners,
                               const float* bmin, const float* bmax, const uint32_t* res,
                               float eps, int lattice, int lag, int want_normals)
{
    ora_mesh_t* m = (ora_mesh_t*)calloc(1, sizeof *m);
    const uint64_t nxy = (uint64_t)res[0] * res[1];
    uint64_t cap = 1024, n = 0;
    float* vtx = (float*)malloc(cap * 6 * sizeof(float));
    uint64_t* keys = (uint64_t*)malloc(cap * sizeof(uint64_t));
    uint8_t* sg = (uint8_t*)malloc(cap);
    float* cur = (float*)calloc(nxy ? 8 * nxy : 1, sizeof(float));
    float* prev = (float*)calloc(nxy ? 8 * nxy : 1, sizeof(float)); /* zero-initialised */
    for (uint32_t z = 0; z < res[2]; ++z) {
        run_slice(ops, nops, corners, bmin, bmax, res, z, lattice, eps, want_normals, cur);
        const float* scan = lag ? prev : cur;
        for (uint64_t c = 0; c < nxy; ++c) { /* main.rs:327-344 */
            const float* t = scan + 8 * c;
            if (!(t[3] > 0.0f)) continue;
            if (n == cap) {
                cap *= 2;
                vtx = (float*)realloc(vtx, cap * 6 * sizeof(float));
                keys = (uint64_t*)realloc(keys, cap * sizeof(uint64_t));
                sg = (uint8_t*)realloc(sg, cap);
            }
            memcpy(vtx + 6 * n, t, 3 * sizeof(float));
            memcpy(vtx + 6 * n + 3, t + 4, 3 * sizeof(float));
            keys[n] = key_of(c % res[0], c / res[0], z); /* label */
            sg[n] = (uint8_t)(uint32_t)t[7];
            ++n;
        }
        if (lag) {
            float* tmp = prev;
            prev = cur;
            cur = tmp;
        }
    }
    free(cur);
    free(prev);
    m->nvtx = n;
    m->vtx = vtx;
    m->keys = keys;
    m->signs = sg;
    m->tri = (uint32_t*)malloc((6 * n + 1) * 3 * sizeof(uint32_t));
    m->ntri = triangles_from_list(keys, sg, n, m->tri, &m->invalid_quads);
    if (lag)
        for (uint64_t i = 0; i < n; ++i) keys[i] -= (uint64_t)1 << 32; /* label -> real z */
    return m;
}

/* Mesh a tape.  lattice=0: the reference's per-cell corner coordinates (8 evaluations
 * per cell); lattice=1: shared lattice points min + h*i.  lag: model the readback lag.
 * nthreads > 0 sets the OpenMP thread count. */
ORA_API ora_mesh_t* ora_mesh(const s2m_op* ops, uint32_t nops, const float* bmin,
                             const float* bmax, const uint32_t* res, float eps, int lattice,
                             int lag, int want_normals)
{
    if (!ora_tape_check(ops, nops)) return NULL;
    return mesh_driver(ops, nops, NULL, bmin, bmax, res, eps, lattice, lag, want_normals);
}

/* Contour given per-cell corners (8 per cell, cells in (z,y,x) order) with the given
 * per-axis cell bounds; normals are zero.  Mirrors s2m_contour_cells. */
ORA_API ora_mesh_t* ora_contour_corners(const float* corners, const uint32_t* res,
                                        const float* cell_lo, const float* cell_hi, int lag)
{
    ora_mesh_t* m = (ora_mesh_t*)calloc(1, sizeof *m);
    const uint64_t nxy = (uint64_t)res[0] * res[1];
    uint64_t cap = 1024, n = 0;
    float* vtx = (float*)malloc(cap * 6 * sizeof(float));
    uint64_t* keys = (uint64_t*)malloc(cap * sizeof(uint64_t));
    uint8_t* sg = (uint8_t*)malloc(cap);
    float* cur = (float*)calloc(nxy ? 8 * nxy : 1, sizeof(float));
    float* prev = (float*)calloc(nxy ? 8 * nxy : 1, sizeof(float));
    const float* lox = cell_lo; const float* loy = cell_lo + res[0];
    const float* loz = cell_lo + res[0] + res[1];
    const float* hix = cell_hi; const float* hiy = cell_hi + res[0];
    const float* hiz = cell_hi + res[0] + res[1];
    for (uint32_t z = 0; z < res[2]; ++z) {
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < (int64_t)nxy; ++c) {
            uint32_t x = (uint32_t)(c % res[0]), y = (uint32_t)(c / res[0]);
            float lo[3] = {lox[x], loy[y], loz[z]}, hi[3] = {hix[x], hiy[y], hiz[z]}, p[3];
            const float* v = corners + 8 * ((uint64_t)z * nxy + (uint64_t)c);
            float* t = cur + 8 * c;
            if (ora_cell_vertex(v, lo, hi, p)) {
                t[0] = p[0]; t[1] = p[1]; t[2] = p[2]; t[3] = 1.0f;
                t[4] = t[5] = t[6] = 0.0f; t[7] = (float)ora_cell_signs(v);
            } else {
                memset(t, 0, 8 * sizeof(float));
            }
        }
        const float* scan = lag ? prev : cur;
        for (uint64_t c = 0; c < nxy; ++c) {
            const float* t = scan + 8 * c;
            if (!(t[3] > 0.0f)) continue;
            if (n == cap) {
                cap *= 2;
                vtx = (float*)realloc(vtx, cap * 6 * sizeof(float));
                keys = (uint64_t*)realloc(keys, cap * sizeof(uint64_t));
                sg = (uint8_t*)realloc(sg, cap);
            }
            memcpy(vtx + 6 * n, t, 3 * sizeof(float));
            memset(vtx + 6 * n + 3, 0, 3 * sizeof(float));
            keys[n] = key_of(c % res[0], c / res[0], z);
            sg[n] = (uint8_t)(uint32_t)t[7];
            ++n;
        }
        if (lag) {
            float* tmp = prev;
            prev = cur;
            cur = tmp;
        }
    }
    free(cur);
    free(prev);
    m->nvtx = n;
    m->vtx = vtx;
    m->keys = keys;
    m->signs = sg;
    m->tri = (uint32_t*)malloc((6 * n + 1) * 3 * sizeof(uint32_t));
    m->ntri = triangles_from_list(keys, sg, n, m->tri, &m->invalid_quads);
    if (lag)
        for (uint64_t i = 0; i < n; ++i) keys[i] -= (uint64_t)1 << 32;
    return m;
}

/* Evaluate the tape at the 8 reference corners of every cell of slices [z0, z1):
 * the per-cell corner dump used to feed both contouring implementations. */
ORA_API void ora_cell_corners(const s2m_op* ops, uint32_t nops, const float* bmin,
                              const float* bmax, const uint32_t* res, uint32_t z0, uint32_t z1,
                              int lattice, float* out)
{
    const int64_t nxy = (int64_t)res[0] * res[1];
    for (uint32_t z = z0; z < z1; ++z) {
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < nxy; ++c) {
            uint32_t x = (uint32_t)(c % res[0]), y = (uint32_t)(c / res[0]);
            float lo[3], hi[3];
            ora_cell_bounds(bmin[0], bmax[0], res[0], x, lattice, &lo[0], &hi[0]);
            ora_cell_bounds(bmin[1], bmax[1], res[1], y, lattice, &lo[1], &hi[1]);
            ora_cell_bounds(bmin[2], bmax[2], res[2], z, lattice, &lo[2], &hi[2]);
            float* v = out + 8 * ((uint64_t)(z - z0) * nxy + (uint64_t)c);
            for (int k = 0; k < 8; ++k)
                v[k] = tape_eval_f32(ops, nops, (k & 1) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1],
                                     (k & 4) ? hi[2] : lo[2]);
        }
    }
}

/* Time the reference algorithm (8 corner evaluations, contour and normal per cell)
 * over slices [z0, z1) of a res^3 job; returns cells processed.  Used for the CPU
 * baseline on a bounded sample of slices. */
ORA_API uint64_t ora_bench_slices(const s2m_op* ops, uint32_t nops, const float* bmin,
                                  const float* bmax, const uint32_t* res, float eps, uint32_t z0,
                                  uint32_t z1, uint64_t* active)
{
    const uint64_t nxy = (uint64_t)res[0] * res[1];
    float* tex = (float*)calloc(8 * nxy, sizeof(float));
    uint64_t act = 0;
    for (uint32_t z = z0; z < z1; ++z) {
        run_slice(ops, nops, NULL, bmin, bmax, res, z, 0, eps, 1, tex);
        for (uint64_t c = 0; c < nxy; ++c) act += tex[8 * c + 3] > 0.0f;
    }
    free(tex);
    *active = act;
    return (uint64_t)(z1 - z0) * nxy;
}

/* ---- Rust `{}` formatting of f32 ------------------------------------------------ */
/* Shortest decimal digits that read back (strtof) as x; among the two L-digit
 * neighbours of x (the exact value truncated to L digits, and that plus one unit) the
 * closer one is preferred and an exact tie goes up, as Rust's flt2dec Dragon shortest
 * mode does.  The exact expansion comes from glibc's correctly rounded printf at 120
 * digits (every f32 has a finite expansion of at most 112 significant digits).  The
 * digits are then laid out in fixed notation as flt2dec::digits_to_dec_str with
 * frac_digits = 0: "0.000ddd", "dd.ddd" or "ddd000". */
ORA_API int ora_format_f32(float x, char* out)
{
    char* o = out;
    if (isnan(x)) return sprintf(out, "NaN");
    if (signbit(x)) *o++ = '-';
    float ax = fabsf(x);
    if (isinf(ax)) { strcpy(o, "inf"); return (int)(o - out) + 3; }
    if (ax == 0.0f) { strcpy(o, "0"); return (int)(o - out) + 1; }
    char ex[160];
    snprintf(ex, sizeof ex, "%.120e", (double)ax);
    char E[130];
    int ne = 0;
    char* q = ex;
    for (; *q && *q != 'e'; ++q)
        if (*q >= '0' && *q <= '9') E[ne++] = *q;
    E[ne] = 0;
    const int e10 = atoi(q + 1); /* ax = E[0].E[1..] * 10^e10 */
    char digits[32];
    int ndig = 0, exp10 = 0;
    for (int L = 1; L <= 9 && !ndig; ++L) {
        long long fl = 0;
        for (int i = 0; i < L; ++i) fl = fl * 10 + (E[i] - '0');
        int cmp = 0; /* rest digits vs "5000..." */
        if (E[L] > '5') cmp = 1;
        else if (E[L] < '5') cmp = -1;
        else
            for (int i = L + 1; i < ne; ++i)
                if (E[i] != '0') { cmp = 1; break; }
        long long cand[2];
        if (cmp >= 0) { cand[0] = fl + 1; cand[1] = fl; }
        else { cand[0] = fl; cand[1] = fl + 1; }
        for (int k = 0; k < 2 && !ndig; ++k) {
            char cb[64];
            snprintf(cb, sizeof cb, "%llde%d", cand[k], e10 - (L - 1));
            if (strtof(cb, NULL) != ax) continue;
            int len = snprintf(cb, sizeof cb, "%lld", cand[k]);
            exp10 = e10 - (L - 1) + len; /* value = 0.cb * 10^exp10 */
            while (len > 1 && cb[len - 1] == '0') cb[--len] = 0;
            memcpy(digits, cb, (size_t)len + 1);
            ndig = len;
        }
    }
    if (exp10 <= 0) {
        *o++ = '0';
        *o++ = '.';
        for (int i = 0; i < -exp10; ++i) *o++ = '0';
        memcpy(o, digits, (size_t)ndig);
        o += ndig;
    } else if (exp10 < ndig) {
        memcpy(o, digits, (size_t)exp10);
        o += exp10;
        *o++ = '.';
        memcpy(o, digits + exp10, (size_t)(ndig - exp10));
        o += ndig - exp10;
    } else {
        memcpy(o, digits, (size_t)ndig);
        o += ndig;
        for (int i = ndig; i < exp10; ++i) *o++ = '0';
    }
    *o = 0;
    return (int)(o - out);
}

/* ---- writers -------------------------------------------------------------------- */
/* STLWriter (mesh.rs:8-48) over TriangleMesh::fetch_triangles (mesh.rs:155-180);
 * facet normal Triangle::normal = (v2 - v0) x (v1 - v0) (lib.rs:182-184) with the
 * cross product (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x). */
ORA_API int ora_write_stl(const float* vtx, uint64_t nvtx, const uint32_t* tri, uint64_t ntri,
                          const char* path)
{
    (void)nvtx;
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    char a[64], b[64], c[64];
    fputs("solid\n", f);
    for (uint64_t t = 0; t < ntri; ++t) {
        const float* v0 = vtx + 6 * (uint64_t)tri[3 * t];
        const float* v1 = vtx + 6 * (uint64_t)tri[3 * t + 1];
        const float* v2 = vtx + 6 * (uint64_t)tri[3 * t + 2];
        float ax = v2[0] - v0[0], ay = v2[1] - v0[1], az = v2[2] - v0[2];
        float bx = v1[0] - v0[0], by = v1[1] - v0[1], bz = v1[2] - v0[2];
        float nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
        ora_format_f32(nx, a); ora_format_f32(ny, b); ora_format_f32(nz, c);
        fprintf(f, "facet normal %s %s %s\n", a, b, c);
        fputs("\touter loop\n", f);
        const float* vs[3] = {v0, v1, v2};
        for (int k = 0; k < 3; ++k) {
            ora_format_f32(vs[k][0], a); ora_format_f32(vs[k][1], b); ora_format_f32(vs[k][2], c);
            fprintf(f, "\t\tvertex %s %s %s\n", a, b, c);
        }
        fputs("\tendloop\n", f);
        fputs("endfacet\n", f);
    }
    fputs("endsolid\n", f);
    return fclose(f) == 0 ? 0 : -1;
}

/* PLYWriter (mesh.rs:50-141) as used by write_ply_to_file (mesh.rs:198-210). */
ORA_API int ora_write_ply(const float* vtx, uint64_t nvtx, const uint32_t* tri, uint64_t ntri,
                          const char* path)
{
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    fprintf(f, "ply\nformat ascii 1.0\ncomment written by rust-sdf\n");
    fprintf(f, "element vertex %llu\n", (unsigned long long)nvtx);
    fprintf(f, "property float x\nproperty float y\nproperty float z\n");
    fprintf(f, "property float nx\nproperty float ny\nproperty float nz\n");
    fprintf(f, "element face %llu\n", (unsigned long long)ntri);
    fprintf(f, "property list uchar int vertex_index\nend_header\n");
    char s[6][64];
    for (uint64_t i = 0; i < nvtx; ++i) {
        for (int k = 0; k < 6; ++k) ora_format_f32(vtx[6 * i + k], s[k]);
        fprintf(f, "%s %s %s %s %s %s\n", s[0], s[1], s[2], s[3], s[4], s[5]);
    }
    for (uint64_t t = 0; t < ntri; ++t)
        fprintf(f, "3 %u %u %u\n", tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]);
    return fclose(f) == 0 ? 0 : -1;
}
