/*
 * facet_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see facet_oracle.h).
 *
 * Compiled with -O2 -ffp-contract=off -fno-fast-math so that the float
 * expressions below round exactly as written (x86-64 SSE, FLT_EVAL_METHOD 0).
 * Citations "core:N" refer to models/histream/src/facetrt/facetrtcore.cpp and
 * "sh/X:N" to models/histream/shader/facetrt/X of the reference.
 */
#include "facet_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_SKY 0xffffffffu

/* ------------------------------------------------------------------ */
/* small vector helpers (float, evaluated left to right as in core:220-266) */
static float dot3(const float a[3], const float b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static void cross3(const float a[3], const float b[3], float out[3])
{
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

static void normalize3(const float a[3], float out[3])
{
    const float length = sqrtf(dot3(a, a));
    const float inv = 1.0f / length;
    out[0] = a[0] * inv;
    out[1] = a[1] * inv;
    out[2] = a[2] * inv;
}

/* ------------------------------------------------------------------ */
/* graph helpers */
void or_graph_free(or_graph *g)
{
    if (!g) return;
    free(g->rowptr);
    free(g->col);
    free(g->cnt);
    free(g->den);
    free(g->sky);
    memset(g, 0, sizeof(*g));
}

void or_graph_copy(const or_graph *g, uint32_t *rowptr, uint32_t *col,
                   uint32_t *cnt, uint32_t *den, uint32_t *sky)
{
    if (rowptr) memcpy(rowptr, g->rowptr, sizeof(uint32_t) * (g->S + 1u));
    if (col) memcpy(col, g->col, sizeof(uint32_t) * g->nnz);
    if (cnt) memcpy(cnt, g->cnt, sizeof(uint32_t) * g->nnz);
    if (den) memcpy(den, g->den, sizeof(uint32_t) * g->S);
    if (sky) memcpy(sky, g->sky, sizeof(uint32_t) * g->S);
}

static int graph_alloc(or_graph *g, uint32_t S, uint32_t nnz)
{
    memset(g, 0, sizeof(*g));
    g->S = S;
    g->nnz = nnz;
    g->rowptr = (uint32_t *)calloc((size_t)S + 1u, sizeof(uint32_t));
    g->col = (uint32_t *)calloc(nnz ? nnz : 1u, sizeof(uint32_t));
    g->cnt = (uint32_t *)calloc(nnz ? nnz : 1u, sizeof(uint32_t));
    g->den = (uint32_t *)calloc(S ? S : 1u, sizeof(uint32_t));
    g->sky = (uint32_t *)calloc(S ? S : 1u, sizeof(uint32_t));
    if (!g->rowptr || !g->col || !g->cnt || !g->den || !g->sky) {
        or_graph_free(g);
        return -2;
    }
    return 0;
}

int or_graph_from_arrays(or_graph *g, uint32_t S, uint32_t nnz,
                         const uint32_t *rowptr, const uint32_t *col,
                         const uint32_t *cnt, const uint32_t *den,
                         const uint32_t *sky)
{
    if (graph_alloc(g, S, nnz)) return -2;
    memcpy(g->rowptr, rowptr, sizeof(uint32_t) * (S + 1u));
    memcpy(g->col, col, sizeof(uint32_t) * nnz);
    memcpy(g->cnt, cnt, sizeof(uint32_t) * nnz);
    memcpy(g->den, den, sizeof(uint32_t) * S);
    memcpy(g->sky, sky, sizeof(uint32_t) * S);
    return 0;
}

/* closure = (sky + sum count) / den - 1, max abs over sides (core:1416-1428) */
static float graph_closure(const or_graph *g)
{
    float closureError = 0.0f;
    for (uint32_t s = 0; s < g->S; ++s) {
        uint64_t visible = g->sky[s];
        for (uint32_t e = g->rowptr[s]; e < g->rowptr[s + 1u]; ++e) visible += g->cnt[e];
        if (g->den[s] != 0u) {
            const float closure = (float)visible / (float)g->den[s];
            const float err = fabsf(closure - 1.0f);
            if (err > closureError) closureError = err;
        }
    }
    return closureError;
}

/* Sort each row's (col, cnt) pairs by column (insertion sort on short rows,
 * qsort on long ones). */
typedef struct { uint32_t col, cnt; } pair_t;
static int pair_cmp(const void *a, const void *b)
{
    const uint32_t x = ((const pair_t *)a)->col, y = ((const pair_t *)b)->col;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* ------------------------------------------------------------------ */
/* directions: core:475-489 and core:491-498 */
void or_hemisphere_directions(uint32_t count, float *xyz)
{
    const float goldenAngle = 2.39996322972865332f;
    for (uint32_t i = 0; i < count; ++i) {
        const float y = ((float)i + 0.5f) / (float)count;
        const float radius = sqrtf(fmaxf(0.0f, 1.0f - y * y));
        const float azimuth = goldenAngle * (float)i;
        xyz[3 * i + 0] = radius * cosf(azimuth);
        xyz[3 * i + 1] = y;
        xyz[3 * i + 2] = radius * sinf(azimuth);
    }
}

void or_sun_direction(float zenith_deg, float azimuth_deg, float xyz[3])
{
    const float kPi = 3.14159265358979323846f;
    const float zenith = zenith_deg * kPi / 180.0f;
    const float azimuth = azimuth_deg * kPi / 180.0f;
    xyz[0] = sinf(zenith) * cosf(azimuth);
    xyz[1] = cosf(zenith);
    xyz[2] = sinf(zenith) * sinf(azimuth);
}

/* ------------------------------------------------------------------ */
/* Hapke background: core:546-605 */
static float clampf_(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

static float hapke_h(float mu, float w)                       /* core:546-550 */
{
    const float g = sqrtf(fmaxf(1.0f - w, 1.0e-6f));
    return (1.0f + 2.0f * mu) / (1.0f + 2.0f * mu * g);
}

static float hapke_kernel(float mu0, float mu, float phase, float w)  /* core:552-563 */
{
    const float pi = 3.14159265358979323846f;
    const float surge = 1.0f / (1.0f + tanf(0.5f * phase) / 0.1f);
    const float multiple = hapke_h(mu0, w) * hapke_h(mu, w) - 1.0f;
    return w / (4.0f * pi) * mu0 / fmaxf(mu0 + mu, 1.0e-5f) * (1.0f + surge + multiple);
}

float or_hapke_reflectance(float reflectance, const float incoming[3], const float outgoing[3],
                           const float normal[3], float strength)      /* core:565-587 */
{
    const float pi = 3.14159265358979323846f;
    float n[3], in[3], out[3];
    normalize3(normal, n);
    normalize3(incoming, in);
    normalize3(outgoing, out);
    if (dot3(in, n) < 0.0f) for (int k = 0; k < 3; ++k) in[k] = in[k] * -1.0f;
    if (dot3(out, n) < 0.0f) for (int k = 0; k < 3; ++k) out[k] = out[k] * -1.0f;
    const float mu0 = fmaxf(dot3(n, in), 1.0e-4f);
    const float mu = fmaxf(dot3(n, out), 1.0e-4f);
    const float phase = acosf(clampf_(dot3(in, out), -1.0f, 1.0f));
    const float base = clampf_(reflectance, 0.0f, 0.999f);
    const float w = clampf_(4.0f * base / powf(1.0f + base, 2.0f), 1.0e-5f, 0.999f);
    const float directional = hapke_kernel(mu0, mu, phase, w);
    const float h1 = hapke_h(1.0f, w);
    const float normalIncidence = w / (8.0f * pi) * h1 * h1;
    const float value = clampf_(base * directional / fmaxf(normalIncidence, 1.0e-6f), 0.0f, 1.0f);
    return base + (value - base) * clampf_(strength, 0.0f, 1.0f);
}

static float hapke_ir(float mu, float w)
{
    const float root = sqrtf(1.0f - w);
    return root * (1.0f + 2.0f * mu / 1.5f) / (1.0f + 2.0f * mu / 1.5f * root);
}

float or_hapke_emissivity(float emissivity, float cosine, float strength)  /* core:589-605 */
{
    const float base = clampf_(emissivity, 0.0f, 1.0f);
    const float w = clampf_(1.0f - base * base, 1.0e-5f, 0.999f);
    const float value = clampf_(base * hapke_ir(clampf_(cosine, 1.0e-4f, 1.0f), w) /
                                    fmaxf(hapke_ir(1.0f, w), 1.0e-6f), 0.0f, 1.0f);
    return base + (value - base) * clampf_(strength, 0.0f, 1.0f);
}

/* ------------------------------------------------------------------ */
/* A-buffer rasterizer: core:1212-1335 */
typedef struct { float depth; uint32_t side; } fragment_t;

typedef struct {
    const or_vertex *v;
    uint32_t F;
    uint32_t W, H, L;
    uint32_t periodic;
    float px, pz;
    float center[3];
    float radius;
    uint32_t *counts;
    fragment_t *frags;
} raster_t;

static const int kTiles[20][2] = {
    {-1, 0}, {1, 0}, {0, -1}, {0, 1}, {-1, -1}, {-1, 1}, {1, -1}, {1, 1},
    {-2, 0}, {2, 0}, {0, -2}, {0, 2}, {-2, -1}, {-2, 1}, {2, -1}, {2, 1},
    {-1, -2}, {1, -2}, {-1, 2}, {1, 2}};

static void tile_offset(uint32_t index, int out[2])
{
    if (index == 0u) {
        out[0] = 0;
        out[1] = 0;
        return;
    }
    const uint32_t k = index - 1u < 19u ? index - 1u : 19u;
    out[0] = kTiles[k][0];
    out[1] = kTiles[k][1];
}

static int raster_init(raster_t *r, const or_vertex *v, uint32_t F,
                       const float bmin[3], const float bmax[3], uint32_t size,
                       uint32_t layers, uint32_t periodic, float px, float pz)
{
    memset(r, 0, sizeof(*r));
    r->v = v;
    r->F = F;
    r->W = size;
    r->H = size;
    r->L = layers;
    r->periodic = periodic;
    r->px = px;
    r->pz = pz;
    float emin[3] = {bmin[0], bmin[1], bmin[2]};
    float emax[3] = {bmax[0], bmax[1], bmax[2]};
    for (uint32_t t = 1u; t <= periodic; ++t) {
        int off[2];
        tile_offset(t, off);
        const float ox = (float)off[0] * px;
        const float oz = (float)off[1] * pz;
        emin[0] = fminf(emin[0], bmin[0] + ox);
        emax[0] = fmaxf(emax[0], bmax[0] + ox);
        emin[2] = fminf(emin[2], bmin[2] + oz);
        emax[2] = fmaxf(emax[2], bmax[2] + oz);
    }
    const float diag[3] = {emax[0] - emin[0], emax[1] - emin[1], emax[2] - emin[2]};
    for (int k = 0; k < 3; ++k) r->center[k] = (emin[k] + emax[k]) * 0.5f;
    r->radius = fmaxf(0.5f * sqrtf(dot3(diag, diag)) * 1.001f, 1.0e-4f);
    r->counts = (uint32_t *)calloc((size_t)size * size, sizeof(uint32_t));
    r->frags = (fragment_t *)malloc((size_t)size * size * layers * sizeof(fragment_t));
    if (!r->counts || !r->frags) {
        free(r->counts);
        free(r->frags);
        return -2;
    }
    return 0;
}

static void raster_free(raster_t *r)
{
    free(r->counts);
    free(r->frags);
}

static float edgef(float ax, float ay, float bx, float by, float px, float py)
{
    return (px - ax) * (by - ay) - (py - ay) * (bx - ax);
}

typedef struct { float x, y, depth; } proj_t;

static void raster_triangle(raster_t *r, const proj_t p[3], uint32_t side, uint32_t *overflow)
{
    const float area = edgef(p[0].x, p[0].y, p[1].x, p[1].y, p[2].x, p[2].y);
    if (fabsf(area) < 1.0e-8f) return;
    const float minx = fminf(fminf(p[0].x, p[1].x), p[2].x);
    const float maxx = fmaxf(fmaxf(p[0].x, p[1].x), p[2].x);
    const float miny = fminf(fminf(p[0].y, p[1].y), p[2].y);
    const float maxy = fmaxf(fmaxf(p[0].y, p[1].y), p[2].y);
    int x0 = (int)floorf(minx);
    if (x0 < 0) x0 = 0;
    int x1 = (int)ceilf(maxx);
    if (x1 > (int)r->W - 1) x1 = (int)r->W - 1;
    int y0 = (int)floorf(miny);
    if (y0 < 0) y0 = 0;
    int y1 = (int)ceilf(maxy);
    if (y1 > (int)r->H - 1) y1 = (int)r->H - 1;
    for (int y = y0; y <= y1; ++y) {
        for (int x = x0; x <= x1; ++x) {
            const float px = (float)x + 0.5f;
            const float py = (float)y + 0.5f;
            const float w0 = edgef(p[1].x, p[1].y, p[2].x, p[2].y, px, py);
            const float w1 = edgef(p[2].x, p[2].y, p[0].x, p[0].y, px, py);
            const float w2 = edgef(p[0].x, p[0].y, p[1].x, p[1].y, px, py);
            const int inside = area > 0.0f ? (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f)
                                           : (w0 <= 0.0f && w1 <= 0.0f && w2 <= 0.0f);
            if (!inside) continue;
            const size_t pixel = (size_t)y * r->W + (size_t)x;
            const uint32_t layer = r->counts[pixel]++;
            if (layer >= r->L) {
                ++*overflow;
                continue;
            }
            const float invArea = 1.0f / area;
            fragment_t *f = &r->frags[pixel * r->L + layer];
            f->depth = (w0 * p[0].depth + w1 * p[1].depth + w2 * p[2].depth) * invArea;
            f->side = side;
        }
    }
}

static uint32_t raster_run(raster_t *r, const float dir[3])
{
    memset(r->counts, 0, sizeof(uint32_t) * (size_t)r->W * r->H);
    float view[3];
    normalize3(dir, view);
    const float refz[3] = {0.0f, 0.0f, 1.0f};
    const float refy[3] = {0.0f, 1.0f, 0.0f};
    const float *reference = fabsf(view[2]) < 0.9f ? refz : refy;
    float rc[3], right[3], up[3];
    cross3(reference, view, rc);
    normalize3(rc, right);
    cross3(view, right, up);
    uint32_t overflow = 0;
    for (uint32_t t = 0; t <= r->periodic; ++t) {
        int off[2];
        tile_offset(t, off);
        const float tx = (float)off[0] * r->px;
        const float tz = (float)off[1] * r->pz;
        for (uint32_t f = 0; f < r->F; ++f) {
            proj_t p[3];
            for (uint32_t c = 0; c < 3u; ++c) {
                const or_vertex *vx = &r->v[f * 3u + c];
                const float pos[3] = {vx->position[0] + tx, vx->position[1], vx->position[2] + tz};
                const float rel[3] = {pos[0] - r->center[0], pos[1] - r->center[1], pos[2] - r->center[2]};
                p[c].x = (0.5f + 0.5f * dot3(rel, right) / r->radius) * (float)r->W;
                p[c].y = (0.5f + 0.5f * dot3(rel, up) / r->radius) * (float)r->H;
                p[c].depth = 0.5f - dot3(rel, view) / (2.0f * r->radius);
            }
            const float *n = r->v[f * 3u].normal;
            const uint32_t side = f * 2u + (dot3(n, view) >= 0.0f ? 0u : 1u);
            raster_triangle(r, p, side, &overflow);
        }
    }
    return overflow;
}

/* open-addressing u64 -> u32 map for undirected edge counts */
typedef struct {
    uint64_t *keys;
    uint32_t *vals;
    size_t cap, size;
} edgemap_t;

static uint64_t mix64(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    return k;
}

static int edgemap_init(edgemap_t *m, size_t cap)
{
    m->cap = cap;
    m->size = 0;
    m->keys = (uint64_t *)malloc(cap * sizeof(uint64_t));
    m->vals = (uint32_t *)calloc(cap, sizeof(uint32_t));
    if (!m->keys || !m->vals) return -2;
    memset(m->keys, 0xff, cap * sizeof(uint64_t));
    return 0;
}

static int edgemap_add(edgemap_t *m, uint64_t key)
{
    if ((m->size + 1) * 2 > m->cap) {
        edgemap_t n;
        if (edgemap_init(&n, m->cap * 2)) return -2;
        for (size_t i = 0; i < m->cap; ++i) {
            if (m->keys[i] == ~0ULL) continue;
            size_t slot = mix64(m->keys[i]) & (n.cap - 1);
            while (n.keys[slot] != ~0ULL) slot = (slot + 1) & (n.cap - 1);
            n.keys[slot] = m->keys[i];
            n.vals[slot] = m->vals[i];
            ++n.size;
        }
        free(m->keys);
        free(m->vals);
        *m = n;
    }
    size_t slot = mix64(key) & (m->cap - 1);
    while (m->keys[slot] != ~0ULL && m->keys[slot] != key) slot = (slot + 1) & (m->cap - 1);
    if (m->keys[slot] == ~0ULL) {
        m->keys[slot] = key;
        ++m->size;
    }
    ++m->vals[slot];
    return 0;
}

/* symmetric CSR from undirected edge counts (core:1398-1412), rows sorted */
static int graph_from_edgemap(or_graph *g, const edgemap_t *m, uint32_t S)
{
    uint32_t *rowcount = (uint32_t *)calloc((size_t)S + 1u, sizeof(uint32_t));
    if (!rowcount) return -2;
    uint64_t nnz = 0;
    for (size_t i = 0; i < m->cap; ++i) {
        if (m->keys[i] == ~0ULL) continue;
        ++rowcount[(uint32_t)(m->keys[i] >> 32) + 1u];
        ++rowcount[(uint32_t)m->keys[i] + 1u];
        nnz += 2;
    }
    uint32_t *den = g->den, *sky = g->sky;
    g->den = NULL;
    g->sky = NULL;
    or_graph tmp;
    if (graph_alloc(&tmp, S, (uint32_t)nnz)) {
        free(rowcount);
        return -2;
    }
    free(tmp.den);
    free(tmp.sky);
    tmp.den = den;
    tmp.sky = sky;
    for (uint32_t s = 0; s < S; ++s) rowcount[s + 1u] += rowcount[s];
    memcpy(tmp.rowptr, rowcount, sizeof(uint32_t) * (S + 1u));
    uint32_t *cursor = rowcount;
    for (size_t i = 0; i < m->cap; ++i) {
        if (m->keys[i] == ~0ULL) continue;
        const uint32_t a = (uint32_t)(m->keys[i] >> 32), b = (uint32_t)m->keys[i];
        tmp.col[cursor[a]] = b;
        tmp.cnt[cursor[a]++] = m->vals[i];
        tmp.col[cursor[b]] = a;
        tmp.cnt[cursor[b]++] = m->vals[i];
    }
    free(rowcount);
    pair_t *buf = (pair_t *)malloc(sizeof(pair_t) * (nnz ? nnz : 1));
    if (!buf) {
        or_graph_free(&tmp);
        return -2;
    }
    for (uint32_t s = 0; s < S; ++s) {
        const uint32_t a = tmp.rowptr[s], b = tmp.rowptr[s + 1u];
        for (uint32_t e = a; e < b; ++e) {
            buf[e - a].col = tmp.col[e];
            buf[e - a].cnt = tmp.cnt[e];
        }
        qsort(buf, b - a, sizeof(pair_t), pair_cmp);
        for (uint32_t e = a; e < b; ++e) {
            tmp.col[e] = buf[e - a].col;
            tmp.cnt[e] = buf[e - a].cnt;
        }
    }
    free(buf);
    *g = tmp;
    return 0;
}

int or_raster_graph(const or_vertex *v, uint32_t F, const float bmin[3],
                    const float bmax[3], uint32_t size, uint32_t layers,
                    uint32_t periodic, float periodicX, float periodicZ,
                    const float *dirs, uint32_t ndirs, or_graph *out)
{
    const uint32_t S = F * 2u;
    raster_t r;
    if (raster_init(&r, v, F, bmin, bmax, size, layers, periodic, periodicX, periodicZ)) return -2;
    memset(out, 0, sizeof(*out));
    out->S = S;
    out->den = (uint32_t *)calloc(S, sizeof(uint32_t));
    out->sky = (uint32_t *)calloc(S, sizeof(uint32_t));
    edgemap_t m;
    if (!out->den || !out->sky || edgemap_init(&m, 1u << 16)) {
        raster_free(&r);
        return -2;
    }
    fragment_t *local = (fragment_t *)malloc(sizeof(fragment_t) * layers);
    uint32_t overflow = 0;
    for (uint32_t di = 0; di < ndirs; ++di) {
        overflow += raster_run(&r, &dirs[3 * di]);
        const size_t pixels = (size_t)r.W * r.H;
        for (size_t pixel = 0; pixel < pixels; ++pixel) {
            const uint32_t count = r.counts[pixel] < r.L ? r.counts[pixel] : r.L;
            if (count == 0u) continue;
            memcpy(local, &r.frags[pixel * r.L], sizeof(fragment_t) * count);
            /* depth sort (core:1374-1375).  Stable insertion sort: the order
             * of exactly equal depths is the one thing std::sort leaves open
             * (parity unpinned for exact depth ties). */
            for (uint32_t i = 1; i < count; ++i) {
                const fragment_t val = local[i];
                uint32_t j = i;
                while (j > 0u && local[j - 1u].depth > val.depth) {
                    local[j] = local[j - 1u];
                    --j;
                }
                local[j] = val;
            }
            for (uint32_t i = 0; i < count; ++i) {
                ++out->den[local[i].side];
                ++out->den[local[i].side ^ 1u];
            }
            ++out->sky[local[0].side];
            ++out->sky[local[count - 1u].side ^ 1u];
            for (uint32_t i = 0; i + 1u < count; ++i) {
                const uint32_t a = local[i].side ^ 1u;
                const uint32_t b = local[i + 1u].side;
                if ((a >> 1) == (b >> 1)) continue;
                const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
                if (edgemap_add(&m, ((uint64_t)lo << 32) | hi)) {
                    free(local);
                    raster_free(&r);
                    return -2;
                }
            }
        }
    }
    free(local);
    raster_free(&r);
    const int rc = graph_from_edgemap(out, &m, S);
    free(m.keys);
    free(m.vals);
    if (rc) return rc;
    out->overflow = overflow;
    out->closure = graph_closure(out);
    return overflow != 0u ? -1 : 0;
}

int or_raster_sunlit(const or_vertex *v, uint32_t F, const float bmin[3],
                     const float bmax[3], uint32_t size, uint32_t layers,
                     uint32_t periodic, float periodicX, float periodicZ,
                     const float sun[3], float *fraction)
{
    const uint32_t S = F * 2u;
    raster_t r;
    if (raster_init(&r, v, F, bmin, bmax, size, layers, periodic, periodicX, periodicZ)) return -2;
    const uint32_t overflow = raster_run(&r, sun);
    if (overflow) {
        raster_free(&r);
        return -1;
    }
    uint32_t *total = (uint32_t *)calloc(S, sizeof(uint32_t));
    uint32_t *visible = (uint32_t *)calloc(S, sizeof(uint32_t));
    const size_t pixels = (size_t)r.W * r.H;
    for (size_t pixel = 0; pixel < pixels; ++pixel) {
        const uint32_t count = r.counts[pixel] < r.L ? r.counts[pixel] : r.L;
        if (count == 0u) continue;
        float nearest = FLT_MAX;
        uint32_t nearestSide = 0u;
        for (uint32_t l = 0; l < count; ++l) {
            const fragment_t *f = &r.frags[pixel * r.L + l];
            ++total[f->side];
            if (f->depth < nearest) {
                nearest = f->depth;
                nearestSide = f->side;
            }
        }
        ++visible[nearestSide];
    }
    for (uint32_t s = 0; s < S; ++s)
        fraction[s] = total[s] ? (float)visible[s] / (float)total[s] : 0.0f;
    free(total);
    free(visible);
    raster_free(&r);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Seeded ray graph: counter RNG, ray generator, watertight test.
 * The HIP path implements the same published methods independently:
 *   - PCG "RXS-M-XS" 32-bit output permutation as a counter hash
 *     (O'Neill 2014), keyed per side and per ray;
 *   - uniform point on a triangle by folding the unit square;
 *   - cosine-weighted direction by Malley's method with rejection sampling
 *     of the unit disk (at most 30 tries, then the normal);
 *   - orthonormal basis of Duff et al. 2017 (JCGT 6(1));
 *   - watertight ray/triangle test of Woop, Benthin and Wald 2013
 *     (JCGT 2(1)) with the double-precision fallback for zero edge values.
 * Closest hit: smallest t > tmin, ties to the smaller facet index; the
 * origin facet is skipped.  Hit side: 2f + (dot(d, n_f) > 0). */
static uint32_t pcg_hash(uint32_t x)
{
    const uint32_t state = x * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

static float u01(uint32_t h)
{
    return (float)(h >> 8) * 5.9604644775390625e-8f;
}

static uint32_t side_key(uint32_t seed, uint32_t side)
{
    return pcg_hash(seed ^ pcg_hash(side));
}

static uint32_t ray_rand(uint32_t key, uint32_t ray, uint32_t k)
{
    return pcg_hash(key ^ pcg_hash(ray * 64u + k));
}

static void triangle_point(const or_vertex *v, uint32_t f, float u, float w, float o[3])
{
    if (u + w > 1.0f) {
        u = 1.0f - u;
        w = 1.0f - w;
    }
    const float *a = v[3u * f].position, *b = v[3u * f + 1u].position, *c = v[3u * f + 2u].position;
    for (int i = 0; i < 3; ++i) {
        const float e1 = b[i] - a[i];
        const float e2 = c[i] - a[i];
        o[i] = a[i] + (u * e1 + w * e2);
    }
}

void or_ray_generate(const or_vertex *v, uint32_t side, uint32_t ray,
                     uint32_t seed, float o[3], float d[3])
{
    const uint32_t f = side >> 1;
    const uint32_t key = side_key(seed, side);
    triangle_point(v, f, u01(ray_rand(key, ray, 0u)), u01(ray_rand(key, ray, 1u)), o);
    float x = 0.0f, y = 0.0f, z = 1.0f;
    for (uint32_t k = 2u; k < 62u; k += 2u) {
        const float a = 2.0f * u01(ray_rand(key, ray, k)) - 1.0f;
        const float b = 2.0f * u01(ray_rand(key, ray, k + 1u)) - 1.0f;
        const float r2 = a * a + b * b;
        if (r2 < 1.0f) {
            x = a;
            y = b;
            z = sqrtf(1.0f - r2);
            break;
        }
    }
    const float *nf = v[3u * f].normal;
    const float sgnSide = (side & 1u) ? -1.0f : 1.0f;
    const float n[3] = {sgnSide * nf[0], sgnSide * nf[1], sgnSide * nf[2]};
    const float sgn = copysignf(1.0f, n[2]);
    const float a = -1.0f / (sgn + n[2]);
    const float b = n[0] * n[1] * a;
    const float t1[3] = {1.0f + sgn * n[0] * n[0] * a, sgn * b, -sgn * n[0]};
    const float t2[3] = {b, sgn + n[1] * n[1] * a, -n[1]};
    for (int i = 0; i < 3; ++i) d[i] = (x * t1[i] + y * t2[i]) + z * n[i];
}

typedef struct {
    float o[3];
    float d[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float inv[3];
} wray_t;

static void wray_init(wray_t *r, const float o[3], const float d[3])
{
    memcpy(r->o, o, sizeof(r->o));
    memcpy(r->d, d, sizeof(r->d));
    int kz = 0;
    if (fabsf(d[1]) > fabsf(d[kz])) kz = 1;
    if (fabsf(d[2]) > fabsf(d[kz])) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (d[kz] < 0.0f) {
        const int tmp = kx;
        kx = ky;
        ky = tmp;
    }
    r->kx = kx;
    r->ky = ky;
    r->kz = kz;
    r->Sx = d[kx] / d[kz];
    r->Sy = d[ky] / d[kz];
    r->Sz = 1.0f / d[kz];
    for (int i = 0; i < 3; ++i) {
        float di = d[i];
        if (fabsf(di) < 1.0e-20f) di = copysignf(1.0e-20f, di);
        r->inv[i] = 1.0f / di;
    }
}

/* returns 1 and *t on hit with t > tmin */
static int watertight(const wray_t *r, const float *p0, const float *p1, const float *p2,
                      float tmin, float *tout)
{
    const float A[3] = {p0[0] - r->o[0], p0[1] - r->o[1], p0[2] - r->o[2]};
    const float B[3] = {p1[0] - r->o[0], p1[1] - r->o[1], p1[2] - r->o[2]};
    const float C[3] = {p2[0] - r->o[0], p2[1] - r->o[1], p2[2] - r->o[2]};
    const float Ax = A[r->kx] - r->Sx * A[r->kz];
    const float Ay = A[r->ky] - r->Sy * A[r->kz];
    const float Bx = B[r->kx] - r->Sx * B[r->kz];
    const float By = B[r->ky] - r->Sy * B[r->kz];
    const float Cx = C[r->kx] - r->Sx * C[r->kz];
    const float Cy = C[r->ky] - r->Sy * C[r->kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    const float det = U + V + W;
    if (det == 0.0f) return 0;
    const float Az = r->Sz * A[r->kz];
    const float Bz = r->Sz * B[r->kz];
    const float Cz = r->Sz * C[r->kz];
    const float T = U * Az + V * Bz + W * Cz;
    const float t = T / det;
    if (!(t > tmin)) return 0;
    *tout = t;
    return 1;
}

/* independent median-split BVH with padded boxes */
typedef struct {
    float lo[3], hi[3];
    uint32_t left;   /* internal: left child index; leaf: first prim */
    uint32_t count;  /* 0 = internal (right = left + 1) */
} obvh_node;

typedef struct {
    obvh_node *nodes;
    uint32_t nnodes;
    uint32_t *prims;
    const or_vertex *v;
    float pad;
} obvh_t;

static float tri_centroid(const or_vertex *v, uint32_t f, int axis)
{
    return (v[3u * f].position[axis] + v[3u * f + 1u].position[axis] + v[3u * f + 2u].position[axis]) * (1.0f / 3.0f);
}

static const or_vertex *g_sort_v;
static int g_sort_axis;
static int prim_cmp(const void *a, const void *b)
{
    const float x = tri_centroid(g_sort_v, *(const uint32_t *)a, g_sort_axis);
    const float y = tri_centroid(g_sort_v, *(const uint32_t *)b, g_sort_axis);
    if (x < y) return -1;
    if (x > y) return 1;
    const uint32_t i = *(const uint32_t *)a, j = *(const uint32_t *)b;
    return i < j ? -1 : (i > j ? 1 : 0);
}

static void obvh_bounds(obvh_t *b, obvh_node *n, uint32_t first, uint32_t count)
{
    for (int k = 0; k < 3; ++k) {
        n->lo[k] = FLT_MAX;
        n->hi[k] = -FLT_MAX;
    }
    for (uint32_t i = first; i < first + count; ++i) {
        const uint32_t f = b->prims[i];
        for (uint32_t c = 0; c < 3u; ++c)
            for (int k = 0; k < 3; ++k) {
                const float x = b->v[3u * f + c].position[k];
                if (x < n->lo[k]) n->lo[k] = x;
                if (x > n->hi[k]) n->hi[k] = x;
            }
    }
    for (int k = 0; k < 3; ++k) {
        n->lo[k] -= b->pad;
        n->hi[k] += b->pad;
    }
}

static void obvh_split(obvh_t *b, uint32_t node, uint32_t first, uint32_t count)
{
    obvh_node *n = &b->nodes[node];
    obvh_bounds(b, n, first, count);
    if (count <= 4u) {
        n->left = first;
        n->count = count;
        return;
    }
    int axis = 0;
    float ext = n->hi[0] - n->lo[0];
    for (int k = 1; k < 3; ++k)
        if (n->hi[k] - n->lo[k] > ext) {
            ext = n->hi[k] - n->lo[k];
            axis = k;
        }
    g_sort_v = b->v;
    g_sort_axis = axis;
    qsort(b->prims + first, count, sizeof(uint32_t), prim_cmp);
    const uint32_t half = count / 2u;
    const uint32_t left = b->nnodes;
    b->nnodes += 2u;
    n->left = left;
    n->count = 0u;
    obvh_split(b, left, first, half);
    obvh_split(b, left + 1u, first + half, count - half);
}

static int obvh_build(obvh_t *b, const or_vertex *v, uint32_t F)
{
    b->v = v;
    b->nodes = (obvh_node *)malloc(sizeof(obvh_node) * (2u * F + 2u));
    b->prims = (uint32_t *)malloc(sizeof(uint32_t) * F);
    if (!b->nodes || !b->prims) return -2;
    float m = 1.0f;
    for (uint32_t i = 0; i < 3u * F; ++i)
        for (int k = 0; k < 3; ++k) {
            const float x = fabsf(v[i].position[k]);
            if (x > m) m = x;
        }
    b->pad = 1.0e-5f * m;
    for (uint32_t i = 0; i < F; ++i) b->prims[i] = i;
    b->nnodes = 1u;
    obvh_split(b, 0u, 0u, F);
    return 0;
}

static void obvh_free(obvh_t *b)
{
    free(b->nodes);
    free(b->prims);
}

static int box_hit(const wray_t *r, const obvh_node *n, float tlimit)
{
    float tn = 0.0f, tf = tlimit;
    for (int k = 0; k < 3; ++k) {
        const float a = (n->lo[k] - r->o[k]) * r->inv[k];
        const float b = (n->hi[k] - r->o[k]) * r->inv[k];
        const float lo = a < b ? a : b, hi = a < b ? b : a;
        if (lo > tn) tn = lo;
        if (hi < tf) tf = hi;
    }
    return tn <= tf;
}

/* closest: mode 0; any hit: mode 1.  The search starts from the best hit
   (*tHit, *facetOut) found so far when `resume` is set (periodic tiles). */
static uint32_t trace_from(const or_vertex *v, uint32_t F, const obvh_t *bvh, const wray_t *r,
                           uint32_t skip, float tmin, int anyhit, int resume, float *tHit, uint32_t *facetOut)
{
    float best = resume ? *tHit : FLT_MAX;
    uint32_t bestF = resume ? *facetOut : OR_SKY;
    if (!bvh) {
        for (uint32_t f = 0; f < F; ++f) {
            if (f == skip) continue;
            float t;
            if (watertight(r, v[3u * f].position, v[3u * f + 1u].position, v[3u * f + 2u].position, tmin, &t)) {
                if (t < best || (t == best && f < bestF)) {
                    best = t;
                    bestF = f;
                    if (anyhit) break;
                }
            }
        }
    } else {
        uint32_t stack[128];
        uint32_t sp = 0;
        stack[sp++] = 0u;
        while (sp) {
            const obvh_node *n = &bvh->nodes[stack[--sp]];
            const float limit = best == FLT_MAX ? FLT_MAX : best * 1.0001f + 1.0e-6f;
            if (!box_hit(r, n, limit)) continue;
            if (n->count) {
                for (uint32_t i = n->left; i < n->left + n->count; ++i) {
                    const uint32_t f = bvh->prims[i];
                    if (f == skip) continue;
                    float t;
                    if (watertight(r, v[3u * f].position, v[3u * f + 1u].position, v[3u * f + 2u].position, tmin, &t)) {
                        if (t < best || (t == best && f < bestF)) {
                            best = t;
                            bestF = f;
                        }
                    }
                }
                if (anyhit && bestF != OR_SKY) break;
            } else {
                stack[sp++] = n->left + 1u;
                stack[sp++] = n->left;
            }
        }
    }
    if (tHit) *tHit = best;
    if (facetOut) *facetOut = bestF;
    if (bestF == OR_SKY) return OR_SKY;
    const float *nf = v[3u * bestF].normal;
    const float dd = r->d[0] * nf[0] + r->d[1] * nf[1] + r->d[2] * nf[2];
    return bestF * 2u + (dd > 0.0f ? 1u : 0u);
}

static uint32_t trace(const or_vertex *v, uint32_t F, const obvh_t *bvh, const wray_t *r,
                      uint32_t skip, float tmin, int anyhit, float *tHit, uint32_t *facetOut)
{
    float t = FLT_MAX;
    uint32_t f = OR_SKY;
    const uint32_t side = trace_from(v, F, bvh, r, skip, tmin, anyhit, 0, &t, &f);
    if (tHit) *tHit = t;
    if (facetOut) *facetOut = f;
    return side;
}

/* Periodic copies (sh/visibility.vert:15-35): tile 0 is the scene, tiles
   1..count the neighbour list of tile_offset(); the ray is traced once per
   tile with its origin moved by minus the tile offset (float ops as
   written), the copies of the origin facet are ordinary surfaces, and the
   best (t, facet) carries over the tiles in order. */
static uint32_t trace_periodic(const or_vertex *v, uint32_t F, const obvh_t *bvh, const float o[3],
                               const float d[3], uint32_t skip, float tmin, int anyhit, uint32_t count, float px,
                               float pz, uint32_t *facetOut)
{
    float t = FLT_MAX;
    uint32_t f = OR_SKY, side = OR_SKY;
    for (uint32_t k = 0; k <= count; ++k) {
        if (anyhit && f != OR_SKY) break;
        int off[2];
        tile_offset(k, off);
        const float ok[3] = {o[0] - (float)off[0] * px, o[1], o[2] - (float)off[1] * pz};
        wray_t r;
        wray_init(&r, k ? ok : o, d);
        side = trace_from(v, F, bvh, &r, k ? OR_SKY : skip, tmin, anyhit, 1, &t, &f);
    }
    if (facetOut) *facetOut = f;
    return side;
}

uint32_t or_trace_one(const or_vertex *v, uint32_t F, const float o[3],
                      const float d[3], uint32_t skipFacet, float tmin, float *tHit)
{
    wray_t r;
    wray_init(&r, o, d);
    return trace(v, F, NULL, &r, skipFacet, tmin, 0, tHit, NULL);
}

static int u32_cmp(const void *a, const void *b)
{
    const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

int or_ray_graph(const or_vertex *v, uint32_t F, uint32_t seed, uint32_t R,
                 float tmin, int brute, or_graph *out, uint32_t *hits)
{
    return or_ray_graph_periodic(v, F, seed, R, tmin, brute, 0u, 1.0f, 1.0f, out, hits);
}

int or_ray_graph_periodic(const or_vertex *v, uint32_t F, uint32_t seed, uint32_t R, float tmin, int brute,
                          uint32_t periodic, float px, float pz, or_graph *out, uint32_t *hits)
{
    const uint32_t S = 2u * F;
    obvh_t bvh;
    memset(&bvh, 0, sizeof(bvh));
    if (!brute && obvh_build(&bvh, v, F)) return -2;
    uint32_t *row = (uint32_t *)malloc(sizeof(uint32_t) * R);
    uint32_t *rowlen = (uint32_t *)calloc(S + 1u, sizeof(uint32_t));
    /* first pass: count unique neighbours per side */
    uint32_t *all = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)S * R);
    if (!row || !rowlen || !all) return -2;
    for (uint32_t s = 0; s < S; ++s) {
        for (uint32_t j = 0; j < R; ++j) {
            float o[3], d[3];
            or_ray_generate(v, s, j, seed, o, d);
            all[(size_t)s * R + j] =
                trace_periodic(v, F, brute ? NULL : &bvh, o, d, s >> 1, tmin, 0, periodic, px, pz, NULL);
        }
    }
    if (hits) memcpy(hits, all, sizeof(uint32_t) * (size_t)S * R);
    uint64_t nnz = 0;
    for (uint32_t s = 0; s < S; ++s) {
        memcpy(row, &all[(size_t)s * R], sizeof(uint32_t) * R);
        qsort(row, R, sizeof(uint32_t), u32_cmp);
        uint32_t u = 0;
        for (uint32_t j = 0; j < R; ++j)
            if (row[j] != OR_SKY && (j == 0 || row[j] != row[j - 1])) ++u;
        rowlen[s + 1u] = u;
        nnz += u;
    }
    if (graph_alloc(out, S, (uint32_t)nnz)) return -2;
    for (uint32_t s = 0; s < S; ++s) out->rowptr[s + 1u] = out->rowptr[s] + rowlen[s + 1u];
    for (uint32_t s = 0; s < S; ++s) {
        memcpy(row, &all[(size_t)s * R], sizeof(uint32_t) * R);
        qsort(row, R, sizeof(uint32_t), u32_cmp);
        uint32_t e = out->rowptr[s];
        uint32_t skyc = 0;
        for (uint32_t j = 0; j < R; ++j) {
            if (row[j] == OR_SKY) {
                ++skyc;
                continue;
            }
            if (j == 0 || row[j] != row[j - 1]) {
                out->col[e] = row[j];
                out->cnt[e] = 1u;
                ++e;
            } else {
                ++out->cnt[e - 1u];
            }
        }
        out->sky[s] = skyc;
        out->den[s] = R;
    }
    out->closure = graph_closure(out);
    free(all);
    free(row);
    free(rowlen);
    if (!brute) obvh_free(&bvh);
    return 0;
}

int or_ray_sunlit(const or_vertex *v, uint32_t F, const float sun[3],
                  uint32_t seed, uint32_t K, float tmin, int brute, float *fraction)
{
    return or_ray_sunlit_periodic(v, F, sun, seed, K, tmin, brute, 0u, 1.0f, 1.0f, fraction);
}

int or_ray_sunlit_periodic(const or_vertex *v, uint32_t F, const float sun[3], uint32_t seed, uint32_t K,
                           float tmin, int brute, uint32_t periodic, float px, float pz, float *fraction)
{
    obvh_t bvh;
    memset(&bvh, 0, sizeof(bvh));
    if (!brute && obvh_build(&bvh, v, F)) return -2;
    for (uint32_t s = 0; s < 2u * F; ++s) {
        const uint32_t f = s >> 1;
        const float *nf = v[3u * f].normal;
        const float sg = (s & 1u) ? -1.0f : 1.0f;
        const float ns[3] = {sg * nf[0], sg * nf[1], sg * nf[2]};
        if (!(dot3(ns, sun) > 0.0f)) {
            fraction[s] = 0.0f;
            continue;
        }
        const uint32_t key = side_key(seed ^ 0x5bd1e995u, s);
        uint32_t lit = 0;
        for (uint32_t j = 0; j < K; ++j) {
            float o[3];
            triangle_point(v, f, u01(ray_rand(key, j, 0u)), u01(ray_rand(key, j, 1u)), o);
            uint32_t hitF;
            trace_periodic(v, F, brute ? NULL : &bvh, o, sun, f, tmin, 1, periodic, px, pz, &hitF);
            if (hitF == OR_SKY) ++lit;
        }
        fraction[s] = (float)lit / (float)K;
    }
    if (!brute) obvh_free(&bvh);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Jacobi solve: core:1476-1529 */
static float incoming_one(const or_graph *g, const float *B, uint32_t s, float sky)
{
    const uint32_t den = g->den[s];
    if (den == 0u) return sky;
    double value = (double)g->sky[s] * sky;
    for (uint32_t e = g->rowptr[s]; e < g->rowptr[s + 1u]; ++e) value += (double)g->cnt[e] * B[g->col[e]];
    return (float)(value / den);
}

void or_solve(const or_graph *g, const or_optics *o, float sky, uint32_t iterations,
              float relaxation, float *B, float *maxDeltaOut)
{
    const uint32_t S = g->S;
    float *cur = (float *)calloc(S, sizeof(float));
    float *nxt = (float *)calloc(S, sizeof(float));
    float maxDelta = 0.0f;
    if (iterations < 1u) iterations = 1u;
    if (relaxation < 0.01f) relaxation = 0.01f;
    if (relaxation > 1.0f) relaxation = 1.0f;
    for (uint32_t it = 0; it < iterations; ++it) {
        maxDelta = 0.0f;
        for (uint32_t f = 0; f < S / 2u; ++f) {
            const uint32_t s0 = 2u * f, s1 = s0 + 1u;
            const float h0 = incoming_one(g, cur, s0, sky) + o[s0].directIrradiance;
            const float h1 = incoming_one(g, cur, s1, sky) + o[s1].directIrradiance;
            const float c0 = o[s0].emission + o[s0].reflectance * h0 + o[s0].transmittance * h1;
            const float c1 = o[s1].emission + o[s1].reflectance * h1 + o[s1].transmittance * h0;
            nxt[s0] = cur[s0] + relaxation * (c0 - cur[s0]);
            nxt[s1] = cur[s1] + relaxation * (c1 - cur[s1]);
            const float d0 = fabsf(nxt[s0] - cur[s0]);
            const float d1 = fabsf(nxt[s1] - cur[s1]);
            if (d0 > maxDelta) maxDelta = d0;
            if (d1 > maxDelta) maxDelta = d1;
        }
        float *t = cur;
        cur = nxt;
        nxt = t;
    }
    memcpy(B, cur, sizeof(float) * S);
    if (maxDeltaOut) *maxDeltaOut = maxDelta;
    free(cur);
    free(nxt);
}

uint32_t or_solve_bands(const or_graph *g, uint32_t nb, const float *r, const float *t,
                        const float *e, const float *D, const float *sky, const float *init,
                        uint32_t maxIter, double tolRel, float relaxation, float *B,
                        double *lastRelOut)
{
    const uint32_t S = g->S;
    const size_t n = (size_t)S * nb;
    float *cur = (float *)calloc(n, sizeof(float));
    float *nxt = (float *)calloc(n, sizeof(float));
    if (init) memcpy(cur, init, sizeof(float) * n);
    uint32_t it = 0;
    double lastRel = 0.0;
    while (it < maxIter) {
        double maxRel = 0.0;
        /* facets are independent within a sweep: the threads split them, every
           value is computed as in the single-threaded loop */
#pragma omp parallel reduction(max : maxRel)
        {
        double *in = (double *)malloc(sizeof(double) * 2u * nb);
#pragma omp for schedule(dynamic, 256)
        for (uint32_t f = 0; f < S / 2u; ++f) {
            for (uint32_t k = 0; k < 2u; ++k) {
                const uint32_t s = 2u * f + k;
                double *acc = in + k * nb;
                const uint32_t den = g->den[s];
                for (uint32_t b = 0; b < nb; ++b) acc[b] = den ? (double)g->sky[s] * sky[b] : 0.0;
                for (uint32_t q = g->rowptr[s]; q < g->rowptr[s + 1u]; ++q) {
                    const float *src = &cur[(size_t)g->col[q] * nb];
                    const double w = (double)g->cnt[q];
                    for (uint32_t b = 0; b < nb; ++b) acc[b] += w * src[b];
                }
                for (uint32_t b = 0; b < nb; ++b) acc[b] = den ? acc[b] / den : (double)sky[b];
            }
            for (uint32_t b = 0; b < nb; ++b) {
                const size_t i0 = (size_t)(2u * f) * nb + b, i1 = i0 + nb;
                const double h0 = in[b] + D[i0];
                const double h1 = in[nb + b] + D[i1];
                const double c0 = (double)e[i0] + (double)r[i0] * h0 + (double)t[i0] * h1;
                const double c1 = (double)e[i1] + (double)r[i1] * h1 + (double)t[i1] * h0;
                nxt[i0] = (float)(cur[i0] + relaxation * (c0 - cur[i0]));
                nxt[i1] = (float)(cur[i1] + relaxation * (c1 - cur[i1]));
                const double d0 = fabs((double)nxt[i0] - cur[i0]) / fmax(fabs((double)nxt[i0]), 1e-30);
                const double d1 = fabs((double)nxt[i1] - cur[i1]) / fmax(fabs((double)nxt[i1]), 1e-30);
                if (d0 > maxRel) maxRel = d0;
                if (d1 > maxRel) maxRel = d1;
            }
        }
        free(in);
        }
        float *tmp = cur;
        cur = nxt;
        nxt = tmp;
        ++it;
        lastRel = maxRel;
        if (maxRel <= tolRel) break;
    }
    memcpy(B, cur, sizeof(float) * n);
    if (lastRelOut) *lastRelOut = lastRel;
    free(cur);
    free(nxt);
    return it;
}

void or_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int or_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* core:1531-1559 */
void or_incident(const or_graph *g, const float *B, const or_optics *o, float sky, float *H)
{
    for (uint32_t s = 0; s < g->S; ++s) {
        double diffuse = sky;
        const uint32_t den = g->den[s];
        if (den != 0u) {
            diffuse = (double)g->sky[s] * sky;
            for (uint32_t e = g->rowptr[s]; e < g->rowptr[s + 1u]; ++e) diffuse += (double)g->cnt[e] * B[g->col[e]];
            diffuse /= den;
        }
        const float h = (float)diffuse + o[s].directIrradiance;
        H[s] = h > 0.0f ? h : 0.0f;
    }
}

void or_incident_bands(const or_graph *g, uint32_t nb, const float *B, const float *D,
                       const float *sky, float *H)
{
    double *acc = (double *)malloc(sizeof(double) * nb);
    for (uint32_t s = 0; s < g->S; ++s) {
        const uint32_t den = g->den[s];
        for (uint32_t b = 0; b < nb; ++b) acc[b] = den ? (double)g->sky[s] * sky[b] : (double)sky[b];
        if (den) {
            for (uint32_t q = g->rowptr[s]; q < g->rowptr[s + 1u]; ++q) {
                const float *src = &B[(size_t)g->col[q] * nb];
                for (uint32_t b = 0; b < nb; ++b) acc[b] += (double)g->cnt[q] * src[b];
            }
            for (uint32_t b = 0; b < nb; ++b) acc[b] /= den;
        }
        for (uint32_t b = 0; b < nb; ++b) {
            const float h = (float)acc[b] + D[(size_t)s * nb + b];
            H[(size_t)s * nb + b] = h > 0.0f ? h : 0.0f;
        }
    }
    free(acc);
}

int or_solve_direct(const or_graph *g, const or_optics *o, float sky, double *B)
{
    const uint32_t S = g->S;
    if (S > 6000u) return -1;
    double *M = (double *)calloc((size_t)S * S, sizeof(double));
    double *rhs = (double *)calloc(S, sizeof(double));
    if (!M || !rhs) return -2;
    /* B0 = e0 + r0 (in0 + D0) + t0 (in1 + D1); in_s = a_s + sum w_sn B_n */
    for (uint32_t s = 0; s < S; ++s) M[(size_t)s * S + s] = 1.0;
    for (uint32_t f = 0; f < S / 2u; ++f) {
        for (uint32_t k = 0; k < 2u; ++k) {
            const uint32_t s = 2u * f + k, q = s ^ 1u;
            const double rr = o[s].reflectance, tt = o[s].transmittance;
            double base = o[s].emission;
            const uint32_t src[2] = {s, q};
            const double coef[2] = {rr, tt};
            for (int m = 0; m < 2; ++m) {
                const uint32_t x = src[m];
                const double c = coef[m];
                double a;
                if (g->den[x] == 0u) {
                    a = sky;
                } else {
                    a = (double)g->sky[x] * sky / g->den[x];
                    for (uint32_t e = g->rowptr[x]; e < g->rowptr[x + 1u]; ++e)
                        M[(size_t)s * S + g->col[e]] -= c * (double)g->cnt[e] / g->den[x];
                }
                base += c * (a + o[x].directIrradiance);
            }
            rhs[s] = base;
        }
    }
    /* Gaussian elimination with partial pivoting */
    for (uint32_t k = 0; k < S; ++k) {
        uint32_t piv = k;
        double best = fabs(M[(size_t)k * S + k]);
        for (uint32_t i = k + 1u; i < S; ++i) {
            const double x = fabs(M[(size_t)i * S + k]);
            if (x > best) {
                best = x;
                piv = i;
            }
        }
        if (best == 0.0) {
            free(M);
            free(rhs);
            return -1;
        }
        if (piv != k) {
            for (uint32_t j = 0; j < S; ++j) {
                const double tmp = M[(size_t)k * S + j];
                M[(size_t)k * S + j] = M[(size_t)piv * S + j];
                M[(size_t)piv * S + j] = tmp;
            }
            const double tmp = rhs[k];
            rhs[k] = rhs[piv];
            rhs[piv] = tmp;
        }
        const double inv = 1.0 / M[(size_t)k * S + k];
        for (uint32_t i = k + 1u; i < S; ++i) {
            const double factor = M[(size_t)i * S + k] * inv;
            if (factor == 0.0) continue;
            double *ri = &M[(size_t)i * S];
            const double *rk = &M[(size_t)k * S];
            for (uint32_t j = k; j < S; ++j) ri[j] -= factor * rk[j];
            rhs[i] -= factor * rhs[k];
        }
    }
    for (uint32_t i = S; i-- > 0;) {
        double acc = rhs[i];
        for (uint32_t j = i + 1u; j < S; ++j) acc -= M[(size_t)i * S + j] * B[j];
        B[i] = acc / M[(size_t)i * S + i];
    }
    free(M);
    free(rhs);
    return 0;
}

/* ------------------------------------------------------------------ */
/* FacetEB in FP64: flux helpers core:854-915, node loop core:3567-3846 */
static const double kSigma = 5.670374419e-8;

static double clampd(double x, double lo, double hi)
{
    return x < lo ? lo : (x > hi ? hi : x);
}

double or_sat_vp(double T)
{
    const double c = clampd(T - 273.15, -80.0, 80.0);
    return 6.107 * pow(10.0, 7.5 * c / (237.3 + c));
}

double or_sat_vp_slope(double T)
{
    const double c = clampd(T - 273.15, -80.0, 80.0);
    return or_sat_vp(T) * 2.3026 * 7.5 * 237.3 / ((237.3 + c) * (237.3 + c));
}

double or_latent(double T, double rho, double P, double ea, double ra, double rs)
{
    const double c = T - 273.15;
    const double lam = (2.501 - 0.002361 * c) * 1.0e6;
    const double deficit = 0.622 * fmax(0.0, or_sat_vp(T) - ea) / fmax(100.0, P);
    return rho * lam * deficit / fmax(1.0, ra + rs);
}

double or_latent_slope(double T, double rho, double P, double ea, double ra, double rs)
{
    const double c = T - 273.15;
    const double lam = (2.501 - 0.002361 * c) * 1.0e6;
    const double deficit = 0.622 * fmax(0.0, or_sat_vp(T) - ea) / fmax(100.0, P);
    const double d = rho / fmax(1.0, ra + rs) *
                     (-2361.0 * deficit + lam * 0.622 * or_sat_vp_slope(T) / fmax(100.0, P));
    return fmax(0.0, d);
}

double or_soil_storage(const or_eb_material *m, double dT)
{
    const double inertia = sqrt(fmax(0.0, (double)m->specificHeat * m->density * m->thermalConductivity));
    return 2.0 * inertia / sqrt(3.14159265358979323846 * fmax(1.0, dT));
}

uint32_t or_eb_node(const or_graph *g, uint32_t F, uint32_t leafCount,
                    const uint32_t *material, const or_eb_material *mats,
                    const float *normalY, const float *absorbedSW,
                    const float *incidentSW,
                    const or_eb_params *p, float *T, const float *Tprev,
                    float *Blw, int *convergedOut, float *maxDeltaTOut,
                    float *maxResEqOut)
{
    const uint32_t S = 2u * F;
    float *r = (float *)malloc(sizeof(float) * S);
    float *tt = (float *)malloc(sizeof(float) * S);
    float *e = (float *)malloc(sizeof(float) * S);
    float *D = (float *)calloc(S, sizeof(float));
    float *B = (float *)calloc(S, sizeof(float));
    float *H = (float *)malloc(sizeof(float) * S);
    double *Rn = (double *)malloc(sizeof(double) * S);
    const double Ta = p->airTemperatureK;
    const double ha = 5.8 + 4.1 * sqrt(fmax(0.1, (double)p->windSpeed));
    const double P = fmax(100.0, (double)p->pressure);
    const double rho = P * 100.0 / (287.05 * Ta);
    const double ra = rho * 1005.0 / fmax(0.1, ha);
    const float sky = p->skyLongwave;
    uint32_t iters = 0;
    double maxDT = 0.0, maxEq = 0.0;

    /* LW optics from T (core:1005-1036); ground sides with the Hapke
     * emissivity at |n_y| (core:1025-1027) */
    #define LW_OPTICS()                                                             \
        for (uint32_t s = 0; s < S; ++s) {                                          \
            const or_eb_material *m = &mats[material[s >> 1]];                     \
            const float rr = fminf(fmaxf(m->tirReflectance, 0.0f), 1.0f);           \
            const float t0 = fminf(fmaxf(m->tirTransmittance, 0.0f), 1.0f - rr);    \
            float eps = fmaxf(0.0f, 1.0f - rr - t0);                                \
            if ((s >> 1) >= leafCount && p->backgroundAngularStrength > 0.0f)       \
                eps = or_hapke_emissivity(eps, fabsf(normalY[s >> 1]),              \
                                          p->backgroundAngularStrength);            \
            r[s] = fminf(fmaxf(1.0f - t0 - eps, 0.0f), 1.0f);                       \
            tt[s] = t0;                                                             \
            const double Tc = clampd(T[s], 1.0, 1000.0);                            \
            e[s] = (float)(eps * kSigma * Tc * Tc * Tc * Tc);                       \
        }

    #define LW_SOLVE()                                                              \
        do {                                                                        \
            if (p->lwTolRel < 0.0f) {                                               \
                memset(B, 0, sizeof(float) * S);                                    \
                or_solve_bands(g, 1u, r, tt, e, D, &sky, NULL, p->lwMaxSweeps,      \
                               -1.0, 1.0f, B, NULL);                                \
            } else {                                                                \
                or_solve_bands(g, 1u, r, tt, e, D, &sky, B, p->lwMaxSweeps,         \
                               p->lwTolRel, 1.0f, B, NULL);                         \
            }                                                                       \
            or_incident_bands(g, 1u, B, D, &sky, H);                                \
        } while (0)

    double *res = (double *)malloc(sizeof(double) * S);
    double *der = (double *)malloc(sizeof(double) * S);

    /* fluxes + residual + derivative per side (leaf facets use side 0) */
    #define FLUXES()                                                                \
        do {                                                                        \
            maxEq = 0.0;                                                            \
            for (uint32_t s = 0; s < S; ++s) {                                      \
                const double eps = clampd(1.0 - (double)r[s] - (double)tt[s], 0.0, 1.0); \
                Rn[s] = (double)absorbedSW[s] + eps * H[s] - (double)e[s];          \
            }                                                                       \
            for (uint32_t f = 0; f < F; ++f) {                                      \
                const or_eb_material *m = &mats[material[f]];                      \
                const uint32_t s0 = 2u * f;                                         \
                if (f < leafCount) {                                                \
                    const double Tl = T[s0];                                        \
                    const double net = Rn[s0] + Rn[s0 + 1u];                        \
                    const double Hs = 2.0 * ha * m->convectiveScale * (Tl - Ta);    \
                    double rs = m->surfaceResistance;                               \
                    if (p->vegetationMethod == 1u) {                                \
                        const double par = fmax(0.0, (double)absorbedSW[s0] + absorbedSW[s0 + 1u]); \
                        const double open = 0.25 + 0.75 * par / (par + 180.0);      \
                        const double tc = Tl - 273.15;                              \
                        const double resp = exp(-pow((tc - 25.0) / 18.0, 2.0));     \
                        rs = clampd(m->surfaceResistance / fmax(0.08, open * resp), 40.0, 100000.0); \
                    }                                                               \
                    const double vf = m->vegetation ? 2.0 : 0.0;                    \
                    const double LE = vf * or_latent(Tl, rho, P, p->vaporPressure, ra, rs); \
                    const double LEd = vf * or_latent_slope(Tl, rho, P, p->vaporPressure, ra, rs); \
                    const double Gd = fmax(0.0, (double)m->heatCapacityPerArea) / p->dTime; \
                    const double G = Gd * (Tl - (double)Tprev[s0]);                 \
                    const double eps = clampd(1.0 - (double)r[s0] - (double)tt[s0], 0.0, 1.0); \
                    /* photovoltaic power and its slope (core:3655-3663, 3772-3778) */ \
                    double Ppv = 0.0, Ppvd = 0.0;                                   \
                    if (m->photovoltaic && incidentSW) {                            \
                        const double inc = (double)incidentSW[s0] + (double)m->pvBifaciality * incidentSW[s0 + 1u]; \
                        const double eta = (double)m->pvEta25 * (1.0 + (double)m->pvGamma * (Tl - 298.15)); \
                        const double avail = (double)absorbedSW[s0] + absorbedSW[s0 + 1u]; \
                        const double raw = eta * inc;                               \
                        Ppv = clampd(raw, 0.0, avail);                              \
                        Ppvd = raw > 0.0 && raw < avail ? (double)m->pvEta25 * m->pvGamma * inc : 0.0; \
                    }                                                               \
                    res[s0] = net - Hs - LE - G - Ppv;                              \
                    der[s0] = 8.0 * eps * kSigma * Tl * Tl * Tl + 2.0 * ha * m->convectiveScale + LEd + Gd + Ppvd; \
                    const double eq = fabs(res[s0]) / fmax(1.0, der[s0]);           \
                    if (eq > maxEq) maxEq = eq;                                     \
                } else {                                                            \
                    for (uint32_t k = 0; k < 2u; ++k) {                             \
                        const uint32_t s = s0 + k;                                  \
                        const double Ts = T[s];                                     \
                        const double Hs = ha * (Ts - Ta);                           \
                        const int dry = m->soilMoisture < 0.03f;                    \
                        const double LE = dry ? 0.0 : or_latent(Ts, rho, P, p->vaporPressure, ra, m->surfaceResistance); \
                        const double LEd = dry ? 0.0 : or_latent_slope(Ts, rho, P, p->vaporPressure, ra, m->surfaceResistance); \
                        double G, Gd;                                               \
                        if (m->soilTemperatureMethod == 0) {                        \
                            Gd = 0.0;                                               \
                            G = 0.35 * Rn[s];                                       \
                        } else {                                                    \
                            Gd = or_soil_storage(m, p->dTime) * (m->soilTemperatureMethod == 2 ? 1.5 : 1.0); \
                            G = Gd * (Ts - (double)Tprev[s]);                       \
                        }                                                           \
                        const double eps = clampd(1.0 - (double)r[s] - (double)tt[s], 0.0, 1.0); \
                        res[s] = Rn[s] - Hs - LE - G;                               \
                        der[s] = 4.0 * eps * kSigma * Ts * Ts * Ts + ha + LEd + Gd; \
                        const double eq = fabs(res[s]) / fmax(1.0, der[s]);         \
                        if (eq > maxEq) maxEq = eq;                                 \
                    }                                                               \
                }                                                                   \
            }                                                                       \
        } while (0)

    for (uint32_t c = 0; c < p->maxCoupling; ++c) {
        LW_OPTICS();
        LW_SOLVE();
        FLUXES();
        maxDT = 0.0;
        for (uint32_t f = 0; f < F; ++f) {
            const uint32_t s0 = 2u * f;
            const uint32_t nk = f < leafCount ? 1u : 2u;
            for (uint32_t k = 0; k < nk; ++k) {
                const uint32_t s = s0 + k;
                const double Tc = T[s];
                const double cand = clampd(Tc + res[s] / fmax(1.0, der[s]), 220.0, 360.0);
                const double upd = Tc + (double)p->relaxation * (cand - Tc);
                const double dT = fabs(upd - Tc);
                if (dT > maxDT) maxDT = dT;
                T[s] = (float)upd;
                if (f < leafCount) T[s0 + 1u] = (float)upd;
            }
        }
        iters = c + 1u;
        if (maxDT <= p->tolerance && maxEq <= p->tolerance) break;
    }
    LW_OPTICS();
    LW_SOLVE();
    FLUXES();
    if (Blw) memcpy(Blw, B, sizeof(float) * S);
    if (convergedOut) *convergedOut = (maxDT <= p->tolerance && maxEq <= p->tolerance);
    if (maxDeltaTOut) *maxDeltaTOut = (float)maxDT;
    if (maxResEqOut) *maxResEqOut = (float)maxEq;
    free(r);
    free(tt);
    free(e);
    free(D);
    free(B);
    free(H);
    free(Rn);
    free(res);
    free(der);
    return iters;
    #undef LW_OPTICS
    #undef LW_SOLVE
    #undef FLUXES
}

/* ------------------------------------------------------------------ */
/* MLP IN -> H -> H -> 1 with tanh-form GELU, FP64 */
static double gelu(double x)
{
    const double k = 0.7978845608028654; /* sqrt(2/pi) */
    return 0.5 * x * (1.0 + tanh(k * (x + 0.044715 * x * x * x)));
}

static double gelu_d(double x)
{
    const double k = 0.7978845608028654;
    const double u = k * (x + 0.044715 * x * x * x);
    const double th = tanh(u);
    const double du = k * (1.0 + 3.0 * 0.044715 * x * x);
    return 0.5 * (1.0 + th) + 0.5 * x * (1.0 - th * th) * du;
}

void or_mlp_forward(const float *P, uint32_t in, uint32_t H, const float *x, uint32_t n, double *y)
{
    const float *W1 = P, *b1 = W1 + (size_t)H * in, *W2 = b1 + H, *b2 = W2 + (size_t)H * H, *W3 = b2 + H, *b3 = W3 + H;
    double *h1 = (double *)malloc(sizeof(double) * H);
    double *h2 = (double *)malloc(sizeof(double) * H);
    for (uint32_t s = 0; s < n; ++s) {
        const float *xs = x + (size_t)s * in;
        for (uint32_t j = 0; j < H; ++j) {
            double a = b1[j];
            for (uint32_t i = 0; i < in; ++i) a += (double)W1[(size_t)j * in + i] * xs[i];
            h1[j] = gelu(a);
        }
        for (uint32_t j = 0; j < H; ++j) {
            double a = b2[j];
            for (uint32_t i = 0; i < H; ++i) a += (double)W2[(size_t)j * H + i] * h1[i];
            h2[j] = gelu(a);
        }
        double o = b3[0];
        for (uint32_t i = 0; i < H; ++i) o += (double)W3[i] * h2[i];
        y[s] = o;
    }
    free(h1);
    free(h2);
}

double or_mlp_loss_grad(const float *P, uint32_t in, uint32_t H, const float *x,
                        const float *target, uint32_t n, double *grad)
{
    const float *W1 = P, *b1 = W1 + (size_t)H * in, *W2 = b1 + H, *b2 = W2 + (size_t)H * H, *W3 = b2 + H, *b3 = W3 + H;
    const size_t np = (size_t)H * in + H + (size_t)H * H + H + H + 1u;
    double *gW1 = grad, *gb1 = gW1 + (size_t)H * in, *gW2 = gb1 + H, *gb2 = gW2 + (size_t)H * H, *gW3 = gb2 + H, *gb3 = gW3 + H;
    memset(grad, 0, sizeof(double) * np);
    double *a1 = (double *)malloc(sizeof(double) * H), *h1 = (double *)malloc(sizeof(double) * H);
    double *a2 = (double *)malloc(sizeof(double) * H), *h2 = (double *)malloc(sizeof(double) * H);
    double *d2 = (double *)malloc(sizeof(double) * H), *d1 = (double *)malloc(sizeof(double) * H);
    double loss = 0.0;
    for (uint32_t s = 0; s < n; ++s) {
        const float *xs = x + (size_t)s * in;
        for (uint32_t j = 0; j < H; ++j) {
            double a = b1[j];
            for (uint32_t i = 0; i < in; ++i) a += (double)W1[(size_t)j * in + i] * xs[i];
            a1[j] = a;
            h1[j] = gelu(a);
        }
        for (uint32_t j = 0; j < H; ++j) {
            double a = b2[j];
            for (uint32_t i = 0; i < H; ++i) a += (double)W2[(size_t)j * H + i] * h1[i];
            a2[j] = a;
            h2[j] = gelu(a);
        }
        double o = b3[0];
        for (uint32_t i = 0; i < H; ++i) o += (double)W3[i] * h2[i];
        const double diff = o - target[s];
        loss += diff * diff;
        const double go = 2.0 * diff / n;
        gb3[0] += go;
        for (uint32_t i = 0; i < H; ++i) {
            gW3[i] += go * h2[i];
            d2[i] = go * W3[i] * gelu_d(a2[i]);
        }
        for (uint32_t j = 0; j < H; ++j) {
            gb2[j] += d2[j];
            for (uint32_t i = 0; i < H; ++i) gW2[(size_t)j * H + i] += d2[j] * h1[i];
        }
        for (uint32_t i = 0; i < H; ++i) {
            double a = 0.0;
            for (uint32_t j = 0; j < H; ++j) a += d2[j] * W2[(size_t)j * H + i];
            d1[i] = a * gelu_d(a1[i]);
        }
        for (uint32_t j = 0; j < H; ++j) {
            gb1[j] += d1[j];
            for (uint32_t i = 0; i < in; ++i) gW1[(size_t)j * in + i] += d1[j] * xs[i];
        }
    }
    free(a1);
    free(h1);
    free(a2);
    free(h2);
    free(d1);
    free(d2);
    return loss / n;
}
