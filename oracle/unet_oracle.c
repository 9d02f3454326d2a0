/*
 * CPU oracle of the learned preconditioner's network (TEST INFRASTRUCTURE
 * ONLY: used by tests/ as the checker of the GPU kernels, never by the
 * product).
 *
 * No reference counterpart: the reference has no network (its nearest hook is
 * src/ML.jl:40-56). This file restates, from their definitions, the U-Net the
 * library trains: 3³ convolutions (bias, ReLU; periodic in x and y, zero
 * padding in z), 2×2×2 mean pooling, nearest upsampling, skip concatenation,
 * a 1×1 head, the mean-square loss against the target scaled by the source's
 * 1/RMS, and the exact gradient of that loss. The input gradient of a
 * convolution is the transposed sum over the taps that reach each input
 * voxel (not a flipped kernel), so the library's flipped-kernel construction
 * is checked, not restated.
 *
 * bf16 = 1 rounds to BF16 (round to nearest even) at the points where the
 * library stores BF16: the 3³ weights, every activation and every gradient
 * tensor. Accumulation is in double. bf16 = 0 is exact double arithmetic,
 * in which mode the gradient is checked against finite differences.
 *
 * Parameter vector (FP32), per 3³ layer l = 0..9: W[tap][c][n] (27·cin·cout),
 * then b[n]; then the head's 16 weights and its bias. tap = (dx+1) +
 * 3(dy+1) + 9(dz+1).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int g_bf16 = 1;

static double rb(double x) {
    if (!g_bf16) return x;
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return f; /* inf / nan */
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

typedef struct {
    int nx, ny, nz;
} vol;
static int nvox(vol g) { return g.nx * g.ny * g.nz; }
static int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }
/* neighbour of voxel v at tap offset (dx, dy, dz): -1 outside in z */
static int nbr(vol g, int v, int dx, int dy, int dz) {
    int nxy = g.nx * g.ny, z = v / nxy, r = v - z * nxy, y = r / g.nx, x = r - y * g.nx;
    int zz = z + dz;
    if (zz < 0 || zz >= g.nz) return -1;
    return wrap(x + dx, g.nx) + g.nx * (wrap(y + dy, g.ny) + g.ny * zz);
}

static const int SPEC[10][2] = {{0, 16}, {16, 16}, {16, 32}, {32, 32}, {32, 64},
                                {64, 64}, {96, 32}, {32, 32}, {48, 16}, {16, 16}};

static void layer_offsets(int inc, size_t *woff, size_t *boff, size_t *head, size_t *total) {
    size_t off = 0;
    for (int l = 0; l < 10; l++) {
        int cin = l == 0 ? inc : SPEC[l][0], cout = SPEC[l][1];
        woff[l] = off;
        off += (size_t)27 * cin * cout;
        boff[l] = off;
        off += cout;
    }
    *head = off;
    *total = off + 17;
}

long long orc_unet_num_params(int inc) {
    size_t w[10], b[10], h, t;
    layer_offsets(inc, w, b, &h, &t);
    return (long long)t;
}

/* out[v][oo + n] = relu(b[n] + Σ in[nb][c] W[tap][c][n]); in has cs_in channels (first cin used) */
static void conv_fwd(vol g, const double *in, int cs_in, int cin, const double *W, const double *b, int cout,
                     double *out, int cs_out, int oo) {
    int N = nvox(g);
    for (int v = 0; v < N; v++) {
        for (int n = 0; n < cout; n++) {
            double s = 0;
            for (int tap = 0; tap < 27; tap++) {
                int u = nbr(g, v, tap % 3 - 1, (tap / 3) % 3 - 1, tap / 9 - 1);
                if (u < 0) continue;
                for (int c = 0; c < cin; c++) s += in[(size_t)u * cs_in + c] * W[((size_t)tap * cin + c) * cout + n];
            }
            s += b[n];
            out[(size_t)v * cs_out + oo + n] = rb(s > 0 ? s : 0);
        }
    }
}
/* gin[u][c] = Σ over (v, tap) with nb(v, tap) = u of gout[v][n] W[tap][c][n] */
static void conv_bwd_data(vol g, const double *gout, int cout, const double *W, int cin, double *gin) {
    int N = nvox(g);
    double *acc = calloc((size_t)N * cin, sizeof(double));
    for (int v = 0; v < N; v++)
        for (int tap = 0; tap < 27; tap++) {
            int u = nbr(g, v, tap % 3 - 1, (tap / 3) % 3 - 1, tap / 9 - 1);
            if (u < 0) continue;
            for (int c = 0; c < cin; c++) {
                double s = 0;
                for (int n = 0; n < cout; n++) s += gout[(size_t)v * cout + n] * W[((size_t)tap * cin + c) * cout + n];
                acc[(size_t)u * cin + c] += s;
            }
        }
    for (size_t e = 0; e < (size_t)N * cin; e++) gin[e] = rb(acc[e]);
    free(acc);
}
/* dW[tap][c][n] = Σ_v in[nb(v, tap)][c] g[v][n]; db[n] = Σ_v g[v][n] */
static void conv_bwd_w(vol g, const double *in, int cs_in, int cin, const double *gv, int cout, double *dW,
                       double *db) {
    int N = nvox(g);
    memset(dW, 0, sizeof(double) * 27 * cin * cout);
    memset(db, 0, sizeof(double) * cout);
    for (int v = 0; v < N; v++) {
        for (int n = 0; n < cout; n++) db[n] += gv[(size_t)v * cout + n];
        for (int tap = 0; tap < 27; tap++) {
            int u = nbr(g, v, tap % 3 - 1, (tap / 3) % 3 - 1, tap / 9 - 1);
            if (u < 0) continue;
            for (int c = 0; c < cin; c++) {
                double x = in[(size_t)u * cs_in + c];
                for (int n = 0; n < cout; n++) dW[((size_t)tap * cin + c) * cout + n] += x * gv[(size_t)v * cout + n];
            }
        }
    }
}
static int child(vol gf, int p, int q) {
    int cx = gf.nx / 2, cy = gf.ny / 2;
    int z = p / (cx * cy), r = p - z * cx * cy, y = r / cx, x = r - y * cx;
    return (2 * x + (q & 1)) + gf.nx * ((2 * y + ((q >> 1) & 1)) + gf.ny * (2 * z + (q >> 2)));
}
static int parent(vol gf, int v) {
    int nxy = gf.nx * gf.ny, z = v / nxy, r = v - z * nxy, y = r / gf.nx, x = r - y * gf.nx;
    return x / 2 + (gf.nx / 2) * (y / 2 + (gf.ny / 2) * (z / 2));
}
static void pool(vol gf, const double *in, int cs, int off, int C, double *out) {
    int Nc = nvox(gf) / 8;
    for (int p = 0; p < Nc; p++)
        for (int c = 0; c < C; c++) {
            double s = 0;
            for (int q = 0; q < 8; q++) s += in[(size_t)child(gf, p, q) * cs + off + c];
            out[(size_t)p * C + c] = rb(s / 8);
        }
}
static void up(vol gf, const double *in, int C, double *out, int cs, int off) {
    int N = nvox(gf);
    for (int v = 0; v < N; v++)
        for (int c = 0; c < C; c++) out[(size_t)v * cs + off + c] = in[(size_t)parent(gf, v) * C + c];
}

typedef struct {
    vol g0, g1, g2;
    int inc;
    double *W[10], *b[10], wh[16], bh;
    double *x0, *a0, *cat0, *a8, *a9, *p0, *a2, *cat1, *a6, *a7, *p1, *a4, *a5, *y;
    double s, rms;
} net;

static void net_init(net *m, const float *params, int nx, int ny, int nz, int inc) {
    size_t wo[10], bo[10], ho, tot;
    layer_offsets(inc, wo, bo, &ho, &tot);
    m->g0 = (vol){nx, ny, nz};
    m->g1 = (vol){nx / 2, ny / 2, nz / 2};
    m->g2 = (vol){nx / 4, ny / 4, nz / 4};
    m->inc = inc;
    for (int l = 0; l < 10; l++) {
        int cin = l == 0 ? inc : SPEC[l][0], cout = SPEC[l][1];
        m->W[l] = malloc(sizeof(double) * 27 * cin * cout);
        m->b[l] = malloc(sizeof(double) * cout);
        for (int i = 0; i < 27 * cin * cout; i++) m->W[l][i] = rb(params[wo[l] + i]);
        for (int n = 0; n < cout; n++) m->b[l][n] = params[bo[l] + n];
    }
    for (int c = 0; c < 16; c++) m->wh[c] = params[ho + c];
    m->bh = params[ho + 16];
    size_t n0 = nvox(m->g0), n1 = nvox(m->g1), n2 = nvox(m->g2);
    m->x0 = calloc(n0 * 2, 8); m->a0 = calloc(n0 * 16, 8); m->cat0 = calloc(n0 * 48, 8);
    m->a8 = calloc(n0 * 16, 8); m->a9 = calloc(n0 * 16, 8); m->y = calloc(n0, 8);
    m->p0 = calloc(n1 * 16, 8); m->a2 = calloc(n1 * 32, 8); m->cat1 = calloc(n1 * 96, 8);
    m->a6 = calloc(n1 * 32, 8); m->a7 = calloc(n1 * 32, 8);
    m->p1 = calloc(n2 * 32, 8); m->a4 = calloc(n2 * 64, 8); m->a5 = calloc(n2 * 64, 8);
}
static void net_free(net *m) {
    for (int l = 0; l < 10; l++) { free(m->W[l]); free(m->b[l]); }
    double *p[] = {m->x0, m->a0, m->cat0, m->a8, m->a9, m->y, m->p0, m->a2, m->cat1, m->a6, m->a7, m->p1, m->a4, m->a5};
    for (int i = 0; i < 14; i++) free(p[i]);
}

static void forward(net *m, const double *src, const double *mask) {
    int n0 = nvox(m->g0);
    double ss = 0;
    for (int v = 0; v < n0; v++) ss += src[v] * src[v];
    m->rms = sqrt(ss / n0);
    m->s = m->rms > 0 ? 1 / m->rms : 0;
    for (int v = 0; v < n0; v++) {
        m->x0[(size_t)v * 2] = rb(src[v] * m->s);
        m->x0[(size_t)v * 2 + 1] = (m->inc == 2 && mask) ? rb(mask[v]) : 0;
    }
    conv_fwd(m->g0, m->x0, 2, m->inc, m->W[0], m->b[0], 16, m->a0, 16, 0);
    conv_fwd(m->g0, m->a0, 16, 16, m->W[1], m->b[1], 16, m->cat0, 48, 32);
    pool(m->g0, m->cat0, 48, 32, 16, m->p0);
    conv_fwd(m->g1, m->p0, 16, 16, m->W[2], m->b[2], 32, m->a2, 32, 0);
    conv_fwd(m->g1, m->a2, 32, 32, m->W[3], m->b[3], 32, m->cat1, 96, 64);
    pool(m->g1, m->cat1, 96, 64, 32, m->p1);
    conv_fwd(m->g2, m->p1, 32, 32, m->W[4], m->b[4], 64, m->a4, 64, 0);
    conv_fwd(m->g2, m->a4, 64, 64, m->W[5], m->b[5], 64, m->a5, 64, 0);
    up(m->g1, m->a5, 64, m->cat1, 96, 0);
    conv_fwd(m->g1, m->cat1, 96, 96, m->W[6], m->b[6], 32, m->a6, 32, 0);
    conv_fwd(m->g1, m->a6, 32, 32, m->W[7], m->b[7], 32, m->a7, 32, 0);
    up(m->g0, m->a7, 32, m->cat0, 48, 0);
    conv_fwd(m->g0, m->cat0, 48, 48, m->W[8], m->b[8], 16, m->a8, 16, 0);
    conv_fwd(m->g0, m->a8, 16, 16, m->W[9], m->b[9], 16, m->a9, 16, 0);
    for (int v = 0; v < n0; v++) {
        double y = m->bh;
        for (int c = 0; c < 16; c++) y += m->wh[c] * m->a9[(size_t)v * 16 + c];
        m->y[v] = y;
    }
}

void orc_unet_forward(const float *params, int nx, int ny, int nz, int inc, const double *src, const double *mask,
                      int bf16, double *out) {
    g_bf16 = bf16;
    net m;
    net_init(&m, params, nx, ny, nz, inc);
    forward(&m, src, mask);
    for (int v = 0; v < nx * ny * nz; v++) out[v] = m.y[v] * m.rms;
    net_free(&m);
}

static void relu_mask(double *g, const double *y, int ycs, int yoff, int C, int N) {
    for (int v = 0; v < N; v++)
        for (int c = 0; c < C; c++)
            if (!(y[(size_t)v * ycs + yoff + c] > 0)) g[(size_t)v * C + c] = 0;
}
static void store_layer(double *grad, size_t wo, size_t bo, const double *dW, const double *db, int cin, int cout) {
    for (int i = 0; i < 27 * cin * cout; i++) grad[wo + i] = dW[i];
    for (int n = 0; n < cout; n++) grad[bo + n] = db[n];
}

/* loss = mean((y − t·s)²) and its gradient with respect to every parameter */
void orc_unet_gradients(const float *params, int nx, int ny, int nz, int inc, const double *src, const double *mask,
                        const double *target, int bf16, double *grad, double *loss) {
    g_bf16 = bf16;
    net m;
    net_init(&m, params, nx, ny, nz, inc);
    forward(&m, src, mask);
    size_t wo[10], bo[10], ho, tot;
    layer_offsets(inc, wo, bo, &ho, &tot);
    vol g0 = m.g0, g1 = m.g1, g2 = m.g2;
    int n0 = nvox(g0), n1 = nvox(g1), n2 = nvox(g2);
    double *dW = malloc(sizeof(double) * 27 * 96 * 64), db[96];
    /* head */
    double L = 0, gbh = 0, gwh[16] = {0};
    double *g = calloc((size_t)n0 * 16, 8);
    for (int v = 0; v < n0; v++) {
        double d = m.y[v] - target[v] * m.s;
        L += d * d;
        double dy = 2 * d / n0;
        gbh += dy;
        for (int c = 0; c < 16; c++) {
            double a = m.a9[(size_t)v * 16 + c];
            gwh[c] += dy * a;
            g[(size_t)v * 16 + c] = a > 0 ? rb(dy * m.wh[c]) : 0;
        }
    }
    *loss = L / n0;
    for (int c = 0; c < 16; c++) grad[ho + c] = gwh[c];
    grad[ho + 16] = gbh;
    /* layer 9 */
    double *t16 = calloc((size_t)n0 * 16, 8), *t48 = calloc((size_t)n0 * 48, 8);
    conv_bwd_w(g0, m.a8, 16, 16, g, 16, dW, db);
    store_layer(grad, wo[9], bo[9], dW, db, 16, 16);
    conv_bwd_data(g0, g, 16, m.W[9], 16, t16);
    relu_mask(t16, m.a8, 16, 0, 16, n0); /* g8 */
    /* layer 8 */
    conv_bwd_w(g0, m.cat0, 48, 48, t16, 16, dW, db);
    store_layer(grad, wo[8], bo[8], dW, db, 48, 16);
    conv_bwd_data(g0, t16, 16, m.W[8], 48, t48); /* gcat0 */
    /* up backward into a7's gradient */
    double *h32 = calloc((size_t)n1 * 32, 8), *k32 = calloc((size_t)n1 * 32, 8), *t96 = calloc((size_t)n1 * 96, 8);
    for (int p = 0; p < n1; p++)
        for (int c = 0; c < 32; c++) {
            double s = 0;
            for (int q = 0; q < 8; q++) s += t48[(size_t)child(g0, p, q) * 48 + c];
            h32[(size_t)p * 32 + c] = m.a7[(size_t)p * 32 + c] > 0 ? rb(s) : 0;
        }
    /* layer 7 */
    conv_bwd_w(g1, m.a6, 32, 32, h32, 32, dW, db);
    store_layer(grad, wo[7], bo[7], dW, db, 32, 32);
    conv_bwd_data(g1, h32, 32, m.W[7], 32, k32);
    relu_mask(k32, m.a6, 32, 0, 32, n1); /* g6 */
    /* layer 6 */
    conv_bwd_w(g1, m.cat1, 96, 96, k32, 32, dW, db);
    store_layer(grad, wo[6], bo[6], dW, db, 96, 32);
    conv_bwd_data(g1, k32, 32, m.W[6], 96, t96); /* gcat1 */
    double *h64 = calloc((size_t)n2 * 64, 8), *k64 = calloc((size_t)n2 * 64, 8), *gp1 = calloc((size_t)n2 * 32, 8);
    for (int p = 0; p < n2; p++)
        for (int c = 0; c < 64; c++) {
            double s = 0;
            for (int q = 0; q < 8; q++) s += t96[(size_t)child(g1, p, q) * 96 + c];
            h64[(size_t)p * 64 + c] = m.a5[(size_t)p * 64 + c] > 0 ? rb(s) : 0;
        }
    /* layer 5 */
    conv_bwd_w(g2, m.a4, 64, 64, h64, 64, dW, db);
    store_layer(grad, wo[5], bo[5], dW, db, 64, 64);
    conv_bwd_data(g2, h64, 64, m.W[5], 64, k64);
    relu_mask(k64, m.a4, 64, 0, 64, n2); /* g4 */
    /* layer 4 */
    conv_bwd_w(g2, m.p1, 32, 32, k64, 64, dW, db);
    store_layer(grad, wo[4], bo[4], dW, db, 32, 64);
    conv_bwd_data(g2, k64, 64, m.W[4], 32, gp1);
    /* skip1: pool backward + concatenation slice */
    for (int v = 0; v < n1; v++)
        for (int c = 0; c < 32; c++) {
            double gv = gp1[(size_t)parent(g1, v) * 32 + c] / 8 + t96[(size_t)v * 96 + 64 + c];
            h32[(size_t)v * 32 + c] = m.cat1[(size_t)v * 96 + 64 + c] > 0 ? rb(gv) : 0;
        }
    /* layer 3 */
    conv_bwd_w(g1, m.a2, 32, 32, h32, 32, dW, db);
    store_layer(grad, wo[3], bo[3], dW, db, 32, 32);
    conv_bwd_data(g1, h32, 32, m.W[3], 32, k32);
    relu_mask(k32, m.a2, 32, 0, 32, n1); /* g2 */
    /* layer 2 */
    double *gp0 = calloc((size_t)n1 * 16, 8);
    conv_bwd_w(g1, m.p0, 16, 16, k32, 32, dW, db);
    store_layer(grad, wo[2], bo[2], dW, db, 16, 32);
    conv_bwd_data(g1, k32, 32, m.W[2], 16, gp0);
    /* skip0 */
    for (int v = 0; v < n0; v++)
        for (int c = 0; c < 16; c++) {
            double gv = gp0[(size_t)parent(g0, v) * 16 + c] / 8 + t48[(size_t)v * 48 + 32 + c];
            g[(size_t)v * 16 + c] = m.cat0[(size_t)v * 48 + 32 + c] > 0 ? rb(gv) : 0;
        }
    /* layer 1 */
    conv_bwd_w(g0, m.a0, 16, 16, g, 16, dW, db);
    store_layer(grad, wo[1], bo[1], dW, db, 16, 16);
    conv_bwd_data(g0, g, 16, m.W[1], 16, t16);
    relu_mask(t16, m.a0, 16, 0, 16, n0); /* g0 */
    /* layer 0 */
    conv_bwd_w(g0, m.x0, 2, inc, t16, 16, dW, db);
    store_layer(grad, wo[0], bo[0], dW, db, inc, 16);
    free(dW); free(g); free(t16); free(t48); free(h32); free(k32); free(t96); free(h64); free(k64); free(gp1);
    free(gp0);
    net_free(&m);
}
