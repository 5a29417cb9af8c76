/* FP64 oracle of the constant-density acoustic PML path, for parity checks
 * at full configuration sizes (the numpy oracle in oracle/acoustic.py takes
 * minutes there).
 *
 * TEST INFRASTRUCTURE: loaded only by tests/ (through oracle/fd64.py); never
 * by the product.
 *
 * Same semantics as oracle/acoustic.py, against which it is checked on small
 * cases (tests/test_oracle_fd64.py):
 *   step    seistorch/equations2d/acoustic.py:73-86
 *             y = A*(2/dt^2*u - B*u_prev + c^2*L u), A = 1/(1/dt^2 + d/dt),
 *             B = 1/dt^2 - d/dt, L the zero-padded 5-point Laplacian / h^2
 *             (acoustic.py:5-71);
 *   loop    seistorch/rnn.py:178-212: zero fields; per step the update, the
 *             raw add of w[n] at the shot's source cell (source.py:47-52),
 *             then the receivers of the post-source field (probe.py:42-44);
 *   adjoint exact discrete transpose of that loop (mu = A*lambda):
 *             mu^k = A*(s^k + 2/dt^2*mu^{k+1} - B*mu^{k+2} + L(c^2*mu^{k+1}))
 *             dJ/dc = sum_n 2c*(L u^n)*mu^{n+1}
 *           with s^k = dJ/d rec[k-1] at the receivers (repeated receivers
 *           add).  The forward fields are rebuilt segment by segment from
 *           checkpoints of the forward (every `every` steps), so memory is
 *           about 2*nt/every + every fields.
 *
 * Shots run one after another; each time step is split over rows with
 * OpenMP.  Build: oracle/Makefile (gcc -O2 -fopenmp -shared). */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int64_t nz, nx, n;
  double two_idt2, ih2;
  double *c2, *A, *Bc;
} Grid;

static int grid_init(Grid* g, const double* c, const double* d, int64_t nz, int64_t nx,
                     double dt, double h) {
  g->nz = nz;
  g->nx = nx;
  g->n = nz * nx;
  g->two_idt2 = 2.0 / (dt * dt);
  g->ih2 = 1.0 / (h * h);
  g->c2 = malloc(sizeof(double) * g->n);
  g->A = malloc(sizeof(double) * g->n);
  g->Bc = malloc(sizeof(double) * g->n);
  if (!g->c2 || !g->A || !g->Bc) return 1;
  for (int64_t i = 0; i < g->n; ++i) {
    g->c2[i] = c[i] * c[i];
    g->A[i] = 1.0 / (1.0 / (dt * dt) + d[i] / dt);
    g->Bc[i] = 1.0 / (dt * dt) - d[i] / dt;
  }
  return 0;
}

static void grid_free(Grid* g) {
  free(g->c2);
  free(g->A);
  free(g->Bc);
}

/* lap(u) at (z, x), zero outside the grid (before the 1/h^2 scaling). */
static inline double lap5(const double* u, int64_t z, int64_t x, int64_t nz, int64_t nx) {
  const double* r = u + z * nx;
  double s = -4.0 * r[x];
  if (z > 0) s += r[x - nx];
  if (z + 1 < nz) s += r[x + nx];
  if (x > 0) s += r[x - 1];
  if (x + 1 < nx) s += r[x + 1];
  return s;
}

/* Steps n0 .. n1-1 of one shot.  f[0] = u^{n0}, f[1] = u^{n0-1}, f[2] spare
 * on entry; f[0] = u^{n1}, f[1] = u^{n1-1} on return (the three buffers
 * rotate).  After step n, u^{n+1} goes to keep slot n+1-n0 (keep != NULL)
 * and its receivers to rec row n (rec != NULL, row stride rec_stride). */
static void advance(const Grid* g, double* f[3], const double* w, int64_t n0, int64_t n1,
                    int64_t src, const int64_t* rcell, int64_t nr, double* rec,
                    int64_t rec_stride, double* keep) {
  double *u = f[0], *up = f[1], *y = f[2];
  const int64_t nz = g->nz, nx = g->nx;
#pragma omp parallel
  {
    for (int64_t n = n0; n < n1; ++n) {
#pragma omp for schedule(static)
      for (int64_t z = 0; z < nz; ++z)
        for (int64_t x = 0; x < nx; ++x) {
          const int64_t i = z * nx + x;
          const double lap = lap5(u, z, x, nz, nx) * g->ih2;
          y[i] = g->A[i] * (g->two_idt2 * u[i] - g->Bc[i] * up[i] + g->c2[i] * lap);
        }
#pragma omp single
      {
        y[src] += w[n];
        if (rec)
          for (int64_t r = 0; r < nr; ++r) rec[n * rec_stride + r] = y[rcell[r]];
        if (keep) memcpy(keep + (n + 1 - n0) * g->n, y, sizeof(double) * g->n);
        double* t = up;
        up = u;
        u = y;
        y = t;
      }
    }
  }
  f[0] = u;
  f[1] = up;
  f[2] = y;
}

/* Records of every shot: rec_out (nt, R), shot b's receivers in columns
 * rec_off[b] .. rec_off[b+1]-1; rec_zx (R x 2) and src_zx (B x 2) are grid
 * (z, x).  Returns 0, or 1 if memory ran out. */
int fd64_acoustic_forward(const double* c, const double* d, int64_t nz, int64_t nx, double dt,
                          double h, const double* w, int64_t nt, const int64_t* src_zx, int64_t B,
                          const int64_t* rec_zx, const int64_t* rec_off, double* rec_out) {
  Grid g;
  if (grid_init(&g, c, d, nz, nx, dt, h)) return 1;
  const int64_t R = rec_off[B];
  double* buf = malloc(sizeof(double) * (size_t)(4 * g.n));
  int64_t* rcell = malloc(sizeof(int64_t) * (size_t)(R + 1));
  if (!buf || !rcell) return 1;
  for (int64_t r = 0; r < R; ++r) rcell[r] = rec_zx[2 * r] * nx + rec_zx[2 * r + 1];
  for (int64_t b = 0; b < B; ++b) {
    memset(buf, 0, sizeof(double) * (size_t)(4 * g.n));
    double* f[3] = {buf, buf + g.n, buf + 2 * g.n};
    const int64_t o = rec_off[b];
    advance(&g, f, w, 0, nt, src_zx[2 * b] * nx + src_zx[2 * b + 1], rcell + o,
            rec_off[b + 1] - o, rec_out + o, R, NULL);
  }
  free(buf);
  free(rcell);
  grid_free(&g);
  return 0;
}

/* dJ/dc (nz, nx), summed over shots, of a functional J of the records with
 * dJ/d rec = rec_cot ((nt, R), the layout of fd64_acoustic_forward).
 * every >= 1: forward checkpoint interval.  Returns 0, or 1 if memory ran
 * out. */
int fd64_acoustic_gradient(const double* c, const double* d, int64_t nz, int64_t nx, double dt,
                           double h, const double* w, int64_t nt, const int64_t* src_zx, int64_t B,
                           const int64_t* rec_zx, const int64_t* rec_off, const double* rec_cot,
                           int64_t every, double* grad) {
  Grid g;
  if (every < 1) every = 1;
  if (grid_init(&g, c, d, nz, nx, dt, h)) return 1;
  const int64_t R = rec_off[B];
  const int64_t nseg = (nt + every - 1) / every; /* segment j: u^{j*every} .. */
  double* ck = malloc(sizeof(double) * (size_t)(2 * nseg * g.n));
  double* seg = malloc(sizeof(double) * (size_t)(every * g.n));
  double* buf = malloc(sizeof(double) * (size_t)(3 * g.n));
  double* mu = malloc(sizeof(double) * (size_t)(3 * g.n));
  if (!ck || !seg || !buf || !mu) return 1;
  memset(grad, 0, sizeof(double) * (size_t)g.n);
  for (int64_t b = 0; b < B; ++b) {
    const int64_t src = src_zx[2 * b] * nx + src_zx[2 * b + 1];
    /* forward, keeping (u^m, u^{m-1}) at the start m = j*every of each segment */
    memset(buf, 0, sizeof(double) * (size_t)(3 * g.n));
    double* f[3] = {buf, buf + g.n, buf + 2 * g.n};
    for (int64_t j = 0; j < nseg; ++j) {
      if (j > 0)
        advance(&g, f, w, (j - 1) * every, j * every, src, NULL, 0, NULL, 0, NULL);
      memcpy(ck + 2 * j * g.n, f[0], sizeof(double) * (size_t)g.n);
      memcpy(ck + (2 * j + 1) * g.n, f[1], sizeof(double) * (size_t)g.n);
    }
    /* reverse: mu1 = mu^{k+1}, mu2 = mu^{k+2} */
    double *mu1 = mu, *mu2 = mu + g.n, *mun = mu + 2 * g.n;
    memset(mu, 0, sizeof(double) * (size_t)(3 * g.n));
    const int64_t o = rec_off[b], nr = rec_off[b + 1] - o;
    int64_t have = -1; /* segment held in seg */
    for (int64_t k = nt; k >= 1; --k) {
      const int64_t j = (k - 1) / every, first = j * every;
      if (j != have) {
        /* rebuild u^{first} .. u^{last-1} into seg slots 0 .. last-1-first */
        const int64_t last = first + every < nt ? first + every : nt;
        double* r[3] = {buf, buf + g.n, buf + 2 * g.n};
        memcpy(r[0], ck + 2 * j * g.n, sizeof(double) * (size_t)g.n);
        memcpy(r[1], ck + (2 * j + 1) * g.n, sizeof(double) * (size_t)g.n);
        memcpy(seg, r[0], sizeof(double) * (size_t)g.n);
        if (last - 1 > first)
          advance(&g, r, w, first, last - 1, src, NULL, 0, NULL, 0, seg);
        have = j;
      }
      const double* uk1 = seg + (k - 1 - first) * g.n; /* u^{k-1} */
#pragma omp parallel
      {
        /* lam = 2/dt^2 mu1 - B mu2 + L(c^2 mu1) */
#pragma omp for schedule(static)
        for (int64_t z = 0; z < nz; ++z)
          for (int64_t x = 0; x < nx; ++x) {
            const int64_t i = z * nx + x;
            double s = -4.0 * (g.c2[i] * mu1[i]);
            if (z > 0) s += g.c2[i - nx] * mu1[i - nx];
            if (z + 1 < nz) s += g.c2[i + nx] * mu1[i + nx];
            if (x > 0) s += g.c2[i - 1] * mu1[i - 1];
            if (x + 1 < nx) s += g.c2[i + 1] * mu1[i + 1];
            mun[i] = g.two_idt2 * mu1[i] - g.Bc[i] * mu2[i] + s * g.ih2;
          }
#pragma omp single
        for (int64_t q = 0; q < nr; ++q) {
          const int64_t cell = rec_zx[2 * (o + q)] * nx + rec_zx[2 * (o + q) + 1];
          mun[cell] += rec_cot[(k - 1) * R + o + q];
        }
        /* mu = A lam; dJ/dc += 2c (L u^{k-1}) mu */
#pragma omp for schedule(static)
        for (int64_t z = 0; z < nz; ++z)
          for (int64_t x = 0; x < nx; ++x) {
            const int64_t i = z * nx + x;
            mun[i] *= g.A[i];
            grad[i] += 2.0 * c[i] * (lap5(uk1, z, x, nz, nx) * g.ih2) * mun[i];
          }
      }
      double* t = mu2;
      mu2 = mu1;
      mu1 = mun;
      mun = t;
    }
  }
  free(ck);
  free(seg);
  free(buf);
  free(mu);
  grid_free(&g);
  return 0;
}
