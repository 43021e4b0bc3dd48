/* Per-pixel alpha blending, forward and adjoint, for one REAL type.
 * TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
 * Included twice by blend.c: REAL = IN = float, and REAL = IN = double.
 *
 * Restates msplat/src/alpha_blending.cu:16-110 (forward) and :112-246
 * (adjoint) pixel by pixel, all channels in one walk (the reference's 32
 * channel chunks only split sums that are linear in the channels).
 * Expressions keep the kernels' operation order.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

static REAL FN(gpower)(REAL a, REAL b, REAL c, REAL dx, REAL dy) {
  return (REAL)-0.5 * ((a * dx) * dx + (c * dy) * dy) - (b * dx) * dy;
}

/* Forward over pixel rows [py0, py1).  Writes out [C,H,W], final_T [H,W],
 * ncontrib [H,W] there; returns the number of (pixel, Gaussian) pairs
 * evaluated. */
static long long FN(blend_forward_rows)(const IN *uv, const IN *conic, const IN *opac, const IN *feat, int C,
                                        const int *idx, const int *range, float bg, int W, int H, int py0, int py1,
                                        REAL *out, REAL *final_T, int *ncontrib) {
  const int gx = (W + 15) / 16;
  long long evaluated = 0;
  REAL *F = (REAL *)calloc(C > 0 ? C : 1, sizeof(REAL));
  for (int py = py0; py < py1; ++py) {
    for (int px = 0; px < W; ++px) {
      const int tile = (py / 16) * gx + (px / 16);
      const int start = range[2 * tile], end = range[2 * tile + 1];
      REAL T = 1;
      int contributor = 0, last = 0;
      for (int c = 0; c < C; ++c) F[c] = 0;
      for (int i = start; i < end; ++i) {
        ++contributor;
        ++evaluated;
        const int g = idx[i];
        const REAL dx = (REAL)uv[2 * g] - (REAL)px, dy = (REAL)uv[2 * g + 1] - (REAL)py;
        const REAL power = FN(gpower)(conic[3 * g], conic[3 * g + 1], conic[3 * g + 2], dx, dy);
        if (power > 0) continue;
        REAL alpha = (REAL)opac[g] * EXPF(power);
        if (!(alpha <= (REAL)0.99f)) alpha = (REAL)0.99f; /* fmin(0.99, x), NaN -> 0.99 */
        if (alpha < (REAL)(1.0f / 255.0f)) continue;
        const REAL nT = T * (1 - alpha);
        if (nT < (REAL)0.0001f) break; /* this pixel is done */
        for (int c = 0; c < C; ++c) F[c] += (REAL)feat[(long long)g * C + c] * alpha * T;
        T = nT;
        last = contributor;
      }
      const long long pix = (long long)py * W + px;
      final_T[pix] = T;
      ncontrib[pix] = last;
      for (int c = 0; c < C; ++c) out[(long long)c * H * W + pix] = F[c] + T * (REAL)bg;
    }
  }
  free(F);
  return evaluated;
}

/* Forward over the whole image (one thread). */
long long FN(blend_forward)(const IN *uv, const IN *conic, const IN *opac, const IN *feat,
                            long long n, int C, const int *idx, const int *range, float bg, int W, int H,
                            REAL *out, REAL *final_T, int *ncontrib) {
  (void)n;
  return FN(blend_forward_rows)(uv, conic, opac, feat, C, idx, range, bg, W, H, 0, H, out, final_T, ncontrib);
}

/* The same with tile rows shared among `threads` OpenMP threads (pixels are
 * independent, so the outputs equal the one-thread call's). */
long long FN(blend_forward_mt)(const IN *uv, const IN *conic, const IN *opac, const IN *feat,
                               long long n, int C, const int *idx, const int *range, float bg, int W, int H,
                               REAL *out, REAL *final_T, int *ncontrib, int threads) {
  (void)n;
  const int gy = (H + 15) / 16;
  long long evaluated = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(+ : evaluated)
  for (int ty = 0; ty < gy; ++ty) {
    const int py1 = 16 * ty + 16 < H ? 16 * ty + 16 : H;
    evaluated += FN(blend_forward_rows)(uv, conic, opac, feat, C, idx, range, bg, W, H, 16 * ty, py1, out, final_T,
                                        ncontrib);
  }
  return evaluated;
}

/* Adjoint over pixel rows [py0, py1), accumulated into the gradient arrays
 * per pixel in row-major pixel order. */
static void FN(blend_backward_rows)(const IN *uv, const IN *conic, const IN *opac, const IN *feat, int C,
                                    const int *idx, const int *range, float bg, int W, int H, int py0, int py1,
                                    const REAL *final_T, const int *ncontrib, const REAL *dout, REAL *duv,
                                    REAL *dconic, REAL *dopac, REAL *dfeat) {
  const int gx = (W + 15) / 16;
  REAL *g = (REAL *)calloc(C > 0 ? C : 1, sizeof(REAL));
  REAL *accum = (REAL *)calloc(C > 0 ? C : 1, sizeof(REAL));
  REAL *last_f = (REAL *)calloc(C > 0 ? C : 1, sizeof(REAL));
  for (int py = py0; py < py1; ++py) {
    for (int px = 0; px < W; ++px) {
      const int tile = (py / 16) * gx + (px / 16);
      const int start = range[2 * tile];
      const long long pix = (long long)py * W + px;
      const REAL T_final = final_T[pix];
      const int last = ncontrib[pix];
      REAL bg_dot = 0;
      for (int c = 0; c < C; ++c) {
        g[c] = dout[(long long)c * H * W + pix];
        bg_dot += (REAL)bg * g[c];
        accum[c] = 0;
        last_f[c] = 0;
      }
      REAL T = T_final, last_alpha = 0;
      for (int i = start + last - 1; i >= start; --i) {
        const int gi = idx[i];
        const REAL a = conic[3 * gi], b = conic[3 * gi + 1], c3 = conic[3 * gi + 2], op = opac[gi];
        const REAL dx = (REAL)uv[2 * gi] - (REAL)px, dy = (REAL)uv[2 * gi + 1] - (REAL)py;
        const REAL power = FN(gpower)(a, b, c3, dx, dy);
        if (power > 0) continue;
        const REAL G = EXPF(power);
        REAL alpha = op * G;
        if (!(alpha <= (REAL)0.99f)) alpha = (REAL)0.99f;
        if (alpha < (REAL)(1.0f / 255.0f)) continue;
        T = T / (1 - alpha);
        const REAL dcol = alpha * T;
        REAL dL_dalpha = 0;
        for (int c = 0; c < C; ++c) {
          const REAL f = feat[(long long)gi * C + c];
          accum[c] = last_alpha * last_f[c] + (1 - last_alpha) * accum[c];
          last_f[c] = f;
          dL_dalpha += (f - accum[c]) * g[c];
          dfeat[(long long)gi * C + c] += dcol * g[c];
        }
        dL_dalpha *= T;
        last_alpha = alpha;
        dL_dalpha += (-T_final / (1 - alpha)) * bg_dot;
        const REAL dL_dG = op * dL_dalpha;
        const REAL gdx = ((-G) * dx) * a - (G * dy) * b;
        const REAL gdy = ((-G) * dy) * c3 - (G * dx) * b;
        duv[2 * gi] += dL_dG * gdx;
        duv[2 * gi + 1] += dL_dG * gdy;
        dconic[3 * gi] += (((-(REAL)0.5 * G) * dx) * dx) * dL_dG;
        dconic[3 * gi + 1] += (((-G) * dx) * dy) * dL_dG;
        dconic[3 * gi + 2] += (((-(REAL)0.5 * G) * dy) * dy) * dL_dG;
        dopac[gi] += G * dL_dalpha;
      }
    }
  }
  free(g);
  free(accum);
  free(last_f);
}

/* Adjoint over the whole image (one thread); gradient outputs are zeroed
 * here. */
void FN(blend_backward)(const IN *uv, const IN *conic, const IN *opac, const IN *feat, long long n,
                        int C, const int *idx, const int *range, float bg, int W, int H, const REAL *final_T,
                        const int *ncontrib, const REAL *dout, REAL *duv, REAL *dconic, REAL *dopac,
                        REAL *dfeat) {
  memset(duv, 0, sizeof(REAL) * 2 * n);
  memset(dconic, 0, sizeof(REAL) * 3 * n);
  memset(dopac, 0, sizeof(REAL) * n);
  memset(dfeat, 0, sizeof(REAL) * n * (C > 0 ? C : 0));
  FN(blend_backward_rows)(uv, conic, opac, feat, C, idx, range, bg, W, H, 0, H, final_T, ncontrib, dout, duv,
                          dconic, dopac, dfeat);
}

/* The same with tile rows shared among `threads` OpenMP threads, each adding
 * into gradient arrays of its own, which are then summed in thread order
 * (a different summation order than the one-thread call). */
void FN(blend_backward_mt)(const IN *uv, const IN *conic, const IN *opac, const IN *feat, long long n,
                           int C, const int *idx, const int *range, float bg, int W, int H, const REAL *final_T,
                           const int *ncontrib, const REAL *dout, REAL *duv, REAL *dconic, REAL *dopac,
                           REAL *dfeat, int threads) {
  const long long per = n * (6 + (C > 0 ? C : 0)); /* duv 2, dconic 3, dopac 1, dfeat C per Gaussian */
  REAL *priv = (REAL *)calloc((size_t)(per > 0 ? per : 1) * (size_t)threads, sizeof(REAL));
  const int gy = (H + 15) / 16;
#pragma omp parallel num_threads(threads)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    REAL *b = priv + (size_t)t * (size_t)per;
#pragma omp for schedule(dynamic, 1)
    for (int ty = 0; ty < gy; ++ty) {
      const int py1 = 16 * ty + 16 < H ? 16 * ty + 16 : H;
      FN(blend_backward_rows)(uv, conic, opac, feat, C, idx, range, bg, W, H, 16 * ty, py1, final_T, ncontrib,
                              dout, b, b + 2 * n, b + 5 * n, b + 6 * n);
    }
  }
  memset(duv, 0, sizeof(REAL) * 2 * n);
  memset(dconic, 0, sizeof(REAL) * 3 * n);
  memset(dopac, 0, sizeof(REAL) * n);
  memset(dfeat, 0, sizeof(REAL) * n * (C > 0 ? C : 0));
  for (int t = 0; t < threads; ++t) {
    const REAL *b = priv + (size_t)t * (size_t)per;
#pragma omp parallel for num_threads(threads)
    for (long long i = 0; i < n; ++i) {
      duv[2 * i] += b[2 * i];
      duv[2 * i + 1] += b[2 * i + 1];
      for (int k = 0; k < 3; ++k) dconic[3 * i + k] += b[2 * n + 3 * i + k];
      dopac[i] += b[5 * n + i];
      for (int c = 0; c < C; ++c) dfeat[i * C + c] += b[6 * n + i * C + c];
    }
  }
  free(priv);
}

#undef FN
#undef CAT
#undef CAT2
