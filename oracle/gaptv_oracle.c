/*
 * FP64 C restatement of the PnP-SCI GAP-TV / ADMM-TV path -- TEST
 * INFRASTRUCTURE ONLY (used by tests/ and by bench.py's accuracy check; the
 * product never links it).
 *
 * Same algorithm as oracle/pnp_sci_oracle.py, in double precision, with
 * OpenMP over pixels (projection) and frames (TV) so that full-size problems
 * (2048x2048x20, 60 iterations) finish in seconds on a multi-core host.
 * Sources restated (paths relative to the reference checkout):
 *   PnP_SCI/python/utils.py:10-26            A_, At_
 *   PnP_SCI/python/pnp_sci_algo.py:625-650   GAP (accelerated / plain)
 *   PnP_SCI/python/pnp_sci_algo.py:793-836   ADMM
 *   PnP_SCI/python/joint_pnp_sci_algo.py:633 clip(theta, 0, 1)
 *   PnP_SCI/python/pnp_sci_algo.py:898-899   weight/eta decay (ADMM_TV_rec)
 *   scikit-image 0.18.3 skimage/restoration/_denoise.py:315-393 (Chambolle)
 * It is pinned against the NumPy oracle (itself bit-exact against
 * scikit-image golden vectors) in tests/test_oracle_c.py.
 *
 * Layout: volumes are frame-major (T,H,W), planes (H,W), C order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Chambolle TV of one HxW frame (_denoise.py:341-393, ndim = 2, tau = 1/4),
 * rows split over the OpenMP team.  work: 3*H*W doubles.  Returns the
 * iteration whose image was returned.  Inner loops are branch-free over the
 * interior columns so they vectorise; the sums are reassociated (only the
 * exit test sees that). */
static int tv_frame(const double* restrict img, double* restrict out, int H, int W,
                    double weight, double eps, int n_iter_max, double* restrict work) {
  const size_t N = (size_t)H * W;
  double* restrict p0 = work;
  double* restrict p1 = work + N;
  double* restrict cur = work + 2 * N; /* `out` of the current iteration */
  const double tau = 0.25, tw = tau / weight;
  double e_first = 0.0, e_prev = 0.0;
  int it = 0;
#pragma omp parallel for schedule(static)
  for (int r = 0; r < H; ++r) {
    const size_t o = (size_t)r * W;
    memset(p0 + o, 0, W * sizeof(double));
    memset(p1 + o, 0, W * sizeof(double));
    memcpy(cur + o, img + o, W * sizeof(double));
  }
  while (it < n_iter_max) {
    double esq = 0.0, gsum = 0.0;
    if (it > 0) {
#pragma omp parallel for schedule(static) reduction(+ : esq)
      for (int r = 0; r < H; ++r) {
        const size_t o = (size_t)r * W;
        const double* restrict up = r > 0 ? p0 + o - W : NULL;
        double s = 0.0;
        {
          double d = -(p0[o] + p1[o]);
          if (up) d += up[0];
          cur[o] = img[o] + d;
          s += d * d;
        }
        if (up) {
#pragma omp simd reduction(+ : s)
          for (int c = 1; c < W; ++c) {
            double d = -(p0[o + c] + p1[o + c]) + up[c] + p1[o + c - 1];
            cur[o + c] = img[o + c] + d;
            s += d * d;
          }
        } else {
#pragma omp simd reduction(+ : s)
          for (int c = 1; c < W; ++c) {
            double d = -(p0[o + c] + p1[o + c]) + p1[o + c - 1];
            cur[o + c] = img[o + c] + d;
            s += d * d;
          }
        }
        esq += s;
      }
    }
#pragma omp parallel for schedule(static) reduction(+ : gsum)
    for (int r = 0; r < H; ++r) {
      const size_t o = (size_t)r * W;
      const int last_row = r == H - 1;
      double s = 0.0;
#pragma omp simd reduction(+ : s)
      for (int c = 0; c < W - 1; ++c) {
        const double g0 = last_row ? 0.0 : cur[o + c + W] - cur[o + c];
        const double g1 = cur[o + c + 1] - cur[o + c];
        double nrm = sqrt(g0 * g0 + g1 * g1);
        s += nrm;
        nrm = nrm * tw + 1.0;
        p0[o + c] = (p0[o + c] - tau * g0) / nrm;
        p1[o + c] = (p1[o + c] - tau * g1) / nrm;
      }
      {
        const int c = W - 1;
        const double g0 = last_row ? 0.0 : cur[o + c + W] - cur[o + c];
        double nrm = sqrt(g0 * g0);
        s += nrm;
        nrm = nrm * tw + 1.0;
        p0[o + c] = (p0[o + c] - tau * g0) / nrm;
        p1[o + c] = p1[o + c] / nrm;
      }
      gsum += s;
    }
    double energy = (esq + weight * gsum) / (double)N;
    if (it == 0) {
      e_first = e_prev = energy;
    } else if (fabs(e_prev - energy) < eps * e_first) {
      break;
    } else {
      e_prev = energy;
    }
    ++it;
  }
  memcpy(out, cur, N * sizeof(double));
  return it < n_iter_max ? it : n_iter_max - 1;
}

/* Per-frame TV of a (T,H,W) volume; returned[t] gets the stop iteration. */
void oracle_tv(const double* img, double* out, int H, int W, int T, double weight, double eps,
               int n_iter_max, int* returned) {
  const size_t N = (size_t)H * W;
  double* work = (double*)malloc(3 * N * sizeof(double));
  for (int t = 0; t < T; ++t) {
    int s = tv_frame(img + t * N, out + t * N, H, W, weight, eps, n_iter_max, work);
    if (returned) returned[t] = s;
  }
  free(work);
}

/*
 * method: 0 accelerated GAP, 1 plain GAP, 2 ADMM.  x0 may be NULL (x0 = At(y)).
 * x (T,H,W) receives the result: GAP returns x after the last TV, ADMM the
 * last projection x.  psnr_sq (optional, `iters` doubles) receives
 * sum((truth - x_k)^2) per iteration when truth is given; exits (optional,
 * `iters` ints) the number of frames whose TV stopped on the energy test
 * before tv_iter_max - 1 updates in each iteration.
 */
void oracle_gap_tv(const double* y, const double* phi, const double* phi_sum, const double* x0,
                   double* x, int H, int W, int T, int method, int clip, int iters, double lam,
                   double gamma, double tv_weight, double eps, int tv_iter_max,
                   double weight_decay, double gamma_decay, const double* truth,
                   double* sq_err, int* exits) {
  const size_t N = (size_t)H * W, V = N * (size_t)T;
  double* yb = (double*)malloc(N * sizeof(double));
  double* y1 = (double*)calloc(N, sizeof(double));
  double* v = (double*)malloc(V * sizeof(double));
  double* theta = (double*)malloc(V * sizeof(double));
  double* b = (double*)calloc(V, sizeof(double));
  int* ret = (int*)malloc((size_t)T * sizeof(int));
  /* x = theta = x0 (or At(y)) */
#pragma omp parallel for
  for (long long i = 0; i < (long long)N; ++i)
    for (int t = 0; t < T; ++t) {
      const size_t o = (size_t)t * N + i;
      x[o] = x0 ? x0[o] : y[i] * phi[o];
      theta[o] = x[o];
    }
  double weight = tv_weight, gam = gamma;
  for (int k = 0; k < iters; ++k) {
    if (method == 2) {
      /* yb = A(theta + b); x = (theta + b) + lam*At((y - yb)/(Phi_sum + gamma)) */
#pragma omp parallel for
      for (long long i = 0; i < (long long)N; ++i) {
        double s = 0.0;
        for (int t = 0; t < T; ++t) {
          const size_t o = (size_t)t * N + i;
          s += (theta[o] + b[o]) * phi[o];
        }
        const double r = (y[i] - s) / (phi_sum[i] + gam);
        for (int t = 0; t < T; ++t) {
          const size_t o = (size_t)t * N + i;
          x[o] = (theta[o] + b[o]) + lam * (r * phi[o]);
          v[o] = x[o] - b[o];
        }
      }
      oracle_tv(v, theta, H, W, T, weight, eps, tv_iter_max, ret);
#pragma omp parallel for
      for (long long o = 0; o < (long long)V; ++o) {
        if (clip) theta[o] = theta[o] < 0.0 ? 0.0 : (theta[o] > 1.0 ? 1.0 : theta[o]);
        b[o] = b[o] - (x[o] - theta[o]);
      }
    } else {
#pragma omp parallel for
      for (long long i = 0; i < (long long)N; ++i) {
        double s = 0.0;
        for (int t = 0; t < T; ++t) s += x[(size_t)t * N + i] * phi[(size_t)t * N + i];
        double r;
        if (method == 0) {
          y1[i] = y1[i] + (y[i] - s);
          r = (y1[i] - s) / phi_sum[i];
        } else {
          r = (y[i] - s) / phi_sum[i];
        }
        for (int t = 0; t < T; ++t) {
          const size_t o = (size_t)t * N + i;
          v[o] = x[o] + lam * (r * phi[o]);
        }
      }
      oracle_tv(v, x, H, W, T, weight, eps, tv_iter_max, ret);
    }
    if (exits) {
      int n = 0;
      for (int t = 0; t < T; ++t) n += ret[t] < tv_iter_max - 1;
      exits[k] = n;
    }
    if (truth && sq_err) {
      double s = 0.0;
#pragma omp parallel for reduction(+ : s)
      for (long long o = 0; o < (long long)V; ++o) {
        const double e = truth[o] - x[o];
        s += e * e;
      }
      sq_err[k] = s;
    }
    weight *= weight_decay;
    gam *= gamma_decay;
  }
  free(yb);
  free(y1);
  free(v);
  free(theta);
  free(b);
  free(ret);
}
