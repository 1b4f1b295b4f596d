/*
 * ws_oracle.c -- TEST INFRASTRUCTURE ONLY.  CPU oracle of the learned
 * warm-start network of iskra_amd/csrc/warmstart.hip (no counterpart in the
 * reference: its Poisson solve is direct).  Forward pass, mean-squared-error
 * loss and backward pass of the same encoder-decoder, in FP64, NHWC, with
 * the same layer order and parameter layout ([cout][3x3 tap][cin] per layer,
 * layers enc0a, enc0b, ..., enc(L-1)b, dec(L-2)a, dec(L-2)b, ..., dec0b,
 * out).
 *
 * round = 0: pure FP64 (pinned by finite differences in the tests).
 * round = 1: values are rounded to bfloat16 (round to nearest even) wherever
 *            the GPU stores bfloat16: weights, the input, every activation,
 *            the masked output gradients that feed the weight- and
 *            input-gradient convolutions, and the loss gradient; sums stay
 *            FP64.  GPU results then agree up to FP32-versus-FP64
 *            accumulation differences.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define WSO_IN 8
#define WSO_OUT 16
#define WSO_MAXL 8

/* channels of level l: base * 2^l, held at base * 8 from level 3 on */
static int wso_ch(int base, int l) { return base << (l < 3 ? l : 3); }

static double wso_bf(double x) {
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return x;
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}

typedef struct {
  int L, base, S, B, round;
} WsoNet;

typedef struct {
  int cin, cout, level, relu;
  const double *w; /* [cout][9][cin] */
  double *gw;
} WsoConv;

static int wso_side(const WsoNet *n, int l) { return n->S >> l; }

/* y[n,h,w,co] = sum W[co][tap][ci] x[n,h+dy,w+dx,ci] (+ skip), relu */
static void wso_conv_fwd(const WsoNet *net, const WsoConv *c, const double *x, const double *skip,
                         double *y) {
  const int H = wso_side(net, c->level);
  for (int n = 0; n < net->B; ++n)
    for (int h = 0; h < H; ++h)
      for (int w = 0; w < H; ++w)
        for (int co = 0; co < c->cout; ++co) {
          double s = 0.0;
          for (int t = 0; t < 9; ++t) {
            const int hh = h + t / 3 - 1, ww = w + t % 3 - 1;
            if (hh < 0 || hh >= H || ww < 0 || ww >= H) continue;
            const double *xp = x + (((size_t)n * H + hh) * H + ww) * c->cin;
            const double *wp = c->w + ((size_t)co * 9 + t) * c->cin;
            for (int ci = 0; ci < c->cin; ++ci) s += wp[ci] * xp[ci];
          }
          const size_t o = (((size_t)n * H + h) * H + w) * c->cout + co;
          if (skip) s += skip[o];
          if (c->relu && s < 0.0) s = 0.0;
          y[o] = net->round ? wso_bf(s) : s;
        }
}

/* gx += conv^T(g) ; gw += x (x) g  (g already masked and rounded) */
static void wso_conv_bwd(const WsoNet *net, WsoConv *c, const double *x, const double *g,
                         double *gx) {
  const int H = wso_side(net, c->level);
  for (int n = 0; n < net->B; ++n)
    for (int h = 0; h < H; ++h)
      for (int w = 0; w < H; ++w)
        for (int co = 0; co < c->cout; ++co) {
          const double gv = g[(((size_t)n * H + h) * H + w) * c->cout + co];
          if (gv == 0.0) continue;
          for (int t = 0; t < 9; ++t) {
            const int hh = h + t / 3 - 1, ww = w + t % 3 - 1;
            if (hh < 0 || hh >= H || ww < 0 || ww >= H) continue;
            const size_t xo = (((size_t)n * H + hh) * H + ww) * c->cin;
            const double *wp = c->w + ((size_t)co * 9 + t) * c->cin;
            double *gwp = c->gw + ((size_t)co * 9 + t) * c->cin;
            for (int ci = 0; ci < c->cin; ++ci) {
              gwp[ci] += gv * x[xo + ci];
              if (gx) gx[xo + ci] += gv * wp[ci];
            }
          }
        }
}

static double *wso_alloc(const WsoNet *n, int l, int C) {
  const size_t s = (size_t)n->B * wso_side(n, l) * wso_side(n, l) * C;
  return (double *)calloc(s ? s : 1, sizeof(double));
}
static size_t wso_count(const WsoNet *n, int l, int C) {
  return (size_t)n->B * wso_side(n, l) * wso_side(n, l) * C;
}

static void wso_pool(const WsoNet *net, int l, int C, const double *x, double *p) {
  const int H = wso_side(net, l), H2 = H / 2;
  for (int n = 0; n < net->B; ++n)
    for (int h = 0; h < H2; ++h)
      for (int w = 0; w < H2; ++w)
        for (int c = 0; c < C; ++c) {
          const size_t b = (((size_t)n * H + 2 * h) * H + 2 * w) * C + c;
          const double s =
              (x[b] + x[b + C]) + (x[b + (size_t)H * C] + x[b + (size_t)H * C + C]);
          const double v = 0.25 * s;
          p[(((size_t)n * H2 + h) * H2 + w) * C + c] = net->round ? wso_bf(v) : v;
        }
}

static void wso_pool_bwd(const WsoNet *net, int l, int C, const double *gp, double *gx) {
  const int H = wso_side(net, l), H2 = H / 2;
  for (int n = 0; n < net->B; ++n)
    for (int h = 0; h < H; ++h)
      for (int w = 0; w < H; ++w)
        for (int c = 0; c < C; ++c)
          gx[(((size_t)n * H + h) * H + w) * C + c] +=
              0.25 * gp[(((size_t)n * H2 + h / 2) * H2 + w / 2) * C + c];
}

static void wso_up(const WsoNet *net, int l, int C, const double *x, double *u) {
  const int H = wso_side(net, l), H2 = H / 2;
  for (int n = 0; n < net->B; ++n)
    for (int h = 0; h < H; ++h)
      for (int w = 0; w < H; ++w)
        for (int c = 0; c < C; ++c)
          u[(((size_t)n * H + h) * H + w) * C + c] =
              x[(((size_t)n * H2 + h / 2) * H2 + w / 2) * C + c];
}

static void wso_up_bwd(const WsoNet *net, int l, int C, const double *gu, double *gx) {
  const int H = wso_side(net, l), H2 = H / 2;
  for (int n = 0; n < net->B; ++n)
    for (int h = 0; h < H; ++h)
      for (int w = 0; w < H; ++w)
        for (int c = 0; c < C; ++c)
          gx[(((size_t)n * H2 + h / 2) * H2 + w / 2) * C + c] +=
              gu[(((size_t)n * H + h) * H + w) * C + c];
}

/* gm = G * [Y > 0] (rounded); skip += unrounded */
static void wso_mask(const WsoNet *net, size_t cnt, const double *G, const double *Y, double *gm,
                     double *skip) {
  for (size_t i = 0; i < cnt; ++i) {
    const double g = Y[i] > 0.0 ? G[i] : 0.0;
    gm[i] = net->round ? wso_bf(g) : g;
    if (skip) skip[i] += g;
  }
}

static int64_t wso_nparam(int L, int base) {
  int64_t n = 0;
  for (int l = 0; l < L; ++l) {
    const int c = wso_ch(base, l), cin = l == 0 ? WSO_IN : (wso_ch(base, l - 1));
    n += 9ll * c * cin + 9ll * c * c;
  }
  for (int l = L - 2; l >= 0; --l) {
    const int c = wso_ch(base, l);
    n += 9ll * c * (wso_ch(base, l + 1)) + 9ll * c * c;
  }
  n += 9ll * WSO_OUT * base;
  return n;
}

int64_t orc_ws_param_count(int L, int base) { return wso_nparam(L, base); }

/* sum of the 4 neighbours of interior pixel (h, x) of plane v, neighbours
 * outside the interior (h, x < S-1) counting 0 */
static double wso_nb(const double *v, int S, int h, int x) {
  double s = 0.0;
  if (h > 0) s += v[(size_t)(h - 1) * S + x];
  if (h < S - 2) s += v[(size_t)(h + 1) * S + x];
  if (x > 0) s += v[(size_t)h * S + x - 1];
  if (x < S - 2) s += v[(size_t)h * S + x + 1];
  return s;
}

/* v <- A v restricted to the interior of each of B planes (w: scratch) */
static void wso_apply_A(int S, int B, const double *v, double *out) {
  const size_t S2 = (size_t)S * S;
  for (int b = 0; b < B; ++b)
    for (int h = 0; h < S; ++h)
      for (int w = 0; w < S; ++w) {
        const size_t i = b * S2 + (size_t)h * S + w;
        out[i] = (h < S - 1 && w < S - 1) ? 4.0 * v[i] - wso_nb(v + b * S2, S, h, w) : 0.0;
      }
}

/* Energy objective of one batch: y_0 = gain o, `sweeps` damped Jacobi sweeps
 * y_{k+1} = y_k + om4 (x - A y_k), e = y_sweeps - gain t,
 * loss = sum e.Ae / sum (gain t).A(gain t), and the output gradient
 * G = 2 gain (I - om4 A)^sweeps A e / denominator on interior pixels (A v =
 * 4 v - neighbours, 0 outside the interior).  x = the network input as it
 * enters the first convolution. */
static double wso_energy(int S, int B, double gain, double om4, int sweeps, const double *x,
                         const double *o, const double *t, double *G) {
  const size_t S2 = (size_t)S * S, n = B * S2;
  double *y = (double *)calloc(n, sizeof(double));
  double *ay = (double *)calloc(n, sizeof(double));
  double *gt = (double *)calloc(n, sizeof(double));
  double num = 0.0, den = 0.0;
  for (int b = 0; b < B; ++b)
    for (int h = 0; h < S - 1; ++h)
      for (int w = 0; w < S - 1; ++w) {
        const size_t i = b * S2 + (size_t)h * S + w;
        y[i] = gain * o[i];
        gt[i] = gain * t[i];
      }
  for (int k = 0; k < sweeps; ++k) {
    wso_apply_A(S, B, y, ay);
    for (size_t i = 0; i < n; ++i)
      if (((i % S2) / S) < (size_t)S - 1 && (i % S) < (size_t)S - 1) y[i] += om4 * (x[i] - ay[i]);
  }
  for (size_t i = 0; i < n; ++i) y[i] -= gt[i]; /* e (0 off the interior) */
  wso_apply_A(S, B, gt, ay);
  for (size_t i = 0; i < n; ++i) den += gt[i] * ay[i];
  wso_apply_A(S, B, y, ay);
  for (size_t i = 0; i < n; ++i) num += y[i] * ay[i];
  if (!(den > 0.0)) den = 1.0;
  if (G) {
    /* h = 2 gain A e / den, then h <- h - om4 A h, sweeps times */
    for (size_t i = 0; i < n; ++i) G[i] = 2.0 * gain / den * ay[i];
    for (int k = 0; k < sweeps; ++k) {
      wso_apply_A(S, B, G, gt);
      for (size_t i = 0; i < n; ++i) G[i] -= om4 * gt[i];
    }
  }
  free(y);
  free(ay);
  free(gt);
  return num / den;
}

/* in: B x S x S normalised inputs; target (nullable): B x S x S; out: B x S x S
 * network output o; grads (nullable): parameter gradients of the loss over
 * interior pixels (h, w < S-1): obj 0 the mean squared error of o against
 * the target, obj 1 the energy objective (wso_energy) with gain, om4 and
 * sweeps.
 * Returns the loss (0 without target). */
double orc_ws_run_obj(int L, int base, int S, int B, int round, int obj, double gain, double om4,
                      int sweeps, const double *params, const double *in, const double *target,
                      double *out, double *grads) {
  WsoNet net = {L, base, S, B, round};
  const int64_t np = wso_nparam(L, base);
  double *w = (double *)malloc(np * sizeof(double));
  for (int64_t i = 0; i < np; ++i) w[i] = round ? wso_bf(params[i]) : params[i];
  double *gw = grads;
  if (gw) memset(gw, 0, np * sizeof(double));
  /* layers */
  WsoConv cv[4 * WSO_MAXL];
  int nl = 0;
  int64_t off = 0;
#define ADD(cin_, cout_, lev_, relu_)                                         \
  do {                                                                       \
    cv[nl].cin = cin_;                                                       \
    cv[nl].cout = cout_;                                                     \
    cv[nl].level = lev_;                                                     \
    cv[nl].relu = relu_;                                                     \
    cv[nl].w = w + off;                                                      \
    cv[nl].gw = gw ? gw + off : NULL;                                        \
    off += 9ll * (cin_) * (cout_);                                           \
    ++nl;                                                                    \
  } while (0)
  for (int l = 0; l < L; ++l) {
    ADD(l == 0 ? WSO_IN : wso_ch(base, l - 1), wso_ch(base, l), l, 1);
    ADD(wso_ch(base, l), wso_ch(base, l), l, 1);
  }
  for (int l = L - 2; l >= 0; --l) {
    ADD(wso_ch(base, l + 1), wso_ch(base, l), l, 1);
    ADD(wso_ch(base, l), wso_ch(base, l), l, 1);
  }
  ADD(base, WSO_OUT, 0, 0);
#undef ADD
  /* activations */
  double *A0 = wso_alloc(&net, 0, WSO_IN);
  double *Ea[WSO_MAXL], *E[WSO_MAXL], *P[WSO_MAXL], *U[WSO_MAXL], *Da[WSO_MAXL], *D[WSO_MAXL];
  double *gEa[WSO_MAXL], *gE[WSO_MAXL], *gP[WSO_MAXL], *gU[WSO_MAXL], *gDa[WSO_MAXL],
      *gD[WSO_MAXL];
  for (int l = 0; l < WSO_MAXL; ++l) {
    Ea[l] = E[l] = P[l] = U[l] = Da[l] = D[l] = NULL;
    gEa[l] = gE[l] = gP[l] = gU[l] = gDa[l] = gD[l] = NULL;
  }
  const size_t S2 = (size_t)S * S;
  for (int b = 0; b < B; ++b)
    for (size_t p = 0; p < S2; ++p) {
      const double v = in[b * S2 + p];
      A0[(b * S2 + p) * WSO_IN] = round ? wso_bf(v) : v;
    }
  for (int l = 0; l < L; ++l) {
    const int c = wso_ch(base, l);
    Ea[l] = wso_alloc(&net, l, c);
    E[l] = wso_alloc(&net, l, c);
    gEa[l] = wso_alloc(&net, l, c);
    gE[l] = wso_alloc(&net, l, c);
    if (l > 0) {
      P[l] = wso_alloc(&net, l, wso_ch(base, l - 1));
      gP[l] = wso_alloc(&net, l, wso_ch(base, l - 1));
      wso_pool(&net, l - 1, wso_ch(base, l - 1), E[l - 1], P[l]);
    }
    wso_conv_fwd(&net, &cv[2 * l], l == 0 ? A0 : P[l], NULL, Ea[l]);
    wso_conv_fwd(&net, &cv[2 * l + 1], Ea[l], NULL, E[l]);
  }
  for (int l = L - 2; l >= 0; --l) {
    const int c = wso_ch(base, l), cu = wso_ch(base, l + 1);
    const double *src = (l + 1 == L - 1) ? E[l + 1] : D[l + 1];
    U[l] = wso_alloc(&net, l, cu);
    gU[l] = wso_alloc(&net, l, cu);
    Da[l] = wso_alloc(&net, l, c);
    gDa[l] = wso_alloc(&net, l, c);
    D[l] = wso_alloc(&net, l, c);
    gD[l] = wso_alloc(&net, l, c);
    wso_up(&net, l, cu, src, U[l]);
    const int li = 2 * L + 2 * (L - 2 - l);
    wso_conv_fwd(&net, &cv[li], U[l], E[l], Da[l]);
    wso_conv_fwd(&net, &cv[li + 1], Da[l], NULL, D[l]);
  }
  const double *last = L > 1 ? D[0] : E[0];
  double *gl = L > 1 ? gD[0] : gE[0];
  double *Y = wso_alloc(&net, 0, WSO_OUT);
  {
    WsoNet unr = net;
    unr.round = 0; /* the output plane is stored in FP32 on the GPU, not bf16 */
    wso_conv_fwd(&unr, &cv[nl - 1], last, NULL, Y);
  }
  for (size_t i = 0; i < B * S2; ++i) out[i] = Y[i * WSO_OUT];
  double loss = 0.0;
  if (target) {
    const double cnt = (double)B * (S - 1) * (S - 1);
    double *G = wso_alloc(&net, 0, WSO_OUT);
    if (obj == 1) {
      double *x0 = (double *)malloc(B * S2 * sizeof(double));
      double *g0 = (double *)calloc(B * S2, sizeof(double));
      for (size_t i = 0; i < B * S2; ++i) x0[i] = A0[i * WSO_IN];
      loss = wso_energy(S, B, gain, om4, sweeps, x0, out, target, g0);
      for (size_t i = 0; i < B * S2; ++i) G[i * WSO_OUT] = round ? wso_bf(g0[i]) : g0[i];
      free(x0);
      free(g0);
    } else {
      for (int b = 0; b < B; ++b)
        for (int h = 0; h < S; ++h)
          for (int x = 0; x < S; ++x) {
            const size_t i = (size_t)b * S2 + (size_t)h * S + x;
            double d = 0.0;
            if (h < S - 1 && x < S - 1) d = out[i] - target[i];
            loss += d * d;
            double g = 2.0 * d / cnt;
            if (round) g = wso_bf((double)(float)(2.0f * (float)d * (float)(1.0 / cnt)));
            G[i * WSO_OUT] = g;
          }
      loss /= cnt;
    }
    if (gw) {
      /* backward */
      wso_conv_bwd(&net, &cv[nl - 1], last, G, gl);
      double *gm = NULL;
      size_t gmcap = 0;
#define GM(cnt_)                                                    \
  do {                                                              \
    if ((cnt_) > gmcap) {                                           \
      free(gm);                                                     \
      gmcap = (cnt_);                                               \
      gm = (double *)malloc(gmcap * sizeof(double));                \
    }                                                               \
  } while (0)
      for (int l = 0; l <= L - 2; ++l) {
        const int c = wso_ch(base, l), cu = wso_ch(base, l + 1);
        const int li = 2 * L + 2 * (L - 2 - l);
        const size_t n1 = wso_count(&net, l, c);
        GM(n1);
        wso_mask(&net, n1, gD[l], D[l], gm, NULL);
        wso_conv_bwd(&net, &cv[li + 1], Da[l], gm, gDa[l]);
        wso_mask(&net, n1, gDa[l], Da[l], gm, gE[l]);
        wso_conv_bwd(&net, &cv[li], U[l], gm, gU[l]);
        double *dst = (l + 1 == L - 1) ? gE[l + 1] : gD[l + 1];
        wso_up_bwd(&net, l, cu, gU[l], dst);
      }
      for (int l = L - 1; l >= 0; --l) {
        const int c = wso_ch(base, l);
        const size_t n1 = wso_count(&net, l, c);
        GM(n1);
        wso_mask(&net, n1, gE[l], E[l], gm, NULL);
        wso_conv_bwd(&net, &cv[2 * l + 1], Ea[l], gm, gEa[l]);
        wso_mask(&net, n1, gEa[l], Ea[l], gm, NULL);
        wso_conv_bwd(&net, &cv[2 * l], l == 0 ? A0 : P[l], gm, l > 0 ? gP[l] : NULL);
        if (l > 0) wso_pool_bwd(&net, l - 1, wso_ch(base, l - 1), gP[l], gE[l - 1]);
      }
#undef GM
      free(gm);
    }
    free(G);
  }
  free(Y);
  free(A0);
  for (int l = 0; l < WSO_MAXL; ++l) {
    free(Ea[l]); free(E[l]); free(P[l]); free(U[l]); free(Da[l]); free(D[l]);
    free(gEa[l]); free(gE[l]); free(gP[l]); free(gU[l]); free(gDa[l]); free(gD[l]);
  }
  free(w);
  return loss;
}

double orc_ws_run(int L, int base, int S, int B, int round, const double *params,
                  const double *in, const double *target, double *out, double *grads) {
  return orc_ws_run_obj(L, base, S, B, round, 0, 1.0, 0.0, 1, params, in, target, out, grads);
}
