// PT2-side kernels: S and B (= H0 - E0 S) assembly, V = <X|H|0> builders and
// the E2 / sumc2 reductions of the eight internally contracted subspaces
// (solve_[A-H]_subspace.F90).
//
// Density conventions (all column-major, na active spinors, na2 = na*na):
//   G1 [a + na*b]                          = <E_ab>
//   G1F[a + na*b]                          = <E_ab F>,      F = sum_w eps_w E_ww
//   G2 [(b + na*a) + na2*(c + na*d)]       = <E_ab E_cd>    (bra pair stored swapped:
//                                              G2 = Y1^H Y1 with Y1 = E|0>)
//   G2F                                     = <E_ab E_cd F>
//   G3 [(b + na*a) + na2*((c + na*d) + na2*(e + na*f))] = <E_ab E_cd E_ef>,  G3F likewise.
// Integral blocks are the transformed subspace blocks, column-major over their
// four indices, e.g. the (s,i|s,i) block H[a + ns*(i + ni*(b + ns*j))] = (ai|bj).
#include "dc2_kernels.hpp"

namespace dc2 {
namespace {

struct DV {
  int na;
  const cplx *G1, *G1F, *G2, *G2F, *G3, *G3F;
  __device__ cplx D1(int a, int b) const { return G1[a + na * b]; }
  __device__ cplx D1F(int a, int b) const { return G1F[a + na * b]; }
  __device__ int64_t i2(int a, int b, int c, int d) const {
    return (int64_t)(b + na * a) + (int64_t)na * na * (c + na * d);
  }
  __device__ int64_t i3(int a, int b, int c, int d, int e, int f) const {
    const int64_t na2 = (int64_t)na * na;
    return (int64_t)(b + na * a) + na2 * ((c + na * d) + na2 * (e + na * f));
  }
  __device__ cplx D2(int a, int b, int c, int d) const { return G2[i2(a, b, c, d)]; }
  __device__ cplx D2F(int a, int b, int c, int d) const { return G2F[i2(a, b, c, d)]; }
  __device__ cplx D3(int a, int b, int c, int d, int e, int f) const { return G3[i3(a, b, c, d, e, f)]; }
  __device__ cplx D3F(int a, int b, int c, int d, int e, int f) const { return G3F[i3(a, b, c, d, e, f)]; }
};

inline DV dv(const Dens& d) { return DV{d.na, d.G1, d.G1F, d.G2, d.G2F, d.G3, d.G3F}; }

__device__ inline cplx neg(cplx a) { return mk(-a.x, -a.y); }

// S and B formula values for j >= i (the computed upper triangle of every sXmat/bXmat)
template <char X>
__device__ void sb_formula(const DV& d, const double* e, double e0, const int* pairs, int64_t i, int64_t j,
                           cplx& s, cplx& b) {
  const int na = d.na;
  if (X == 'A') {  // solve_A_subspace.F90:312-330, 374-409
    const int x = (int)(i / (na * na)), y = (int)((i / na) % na), z = (int)(i % na);
    const int t = (int)(j / (na * na)), u = (int)((j / na) % na), v = (int)(j % na);
    s = neg(d.D3(z, y, t, x, u, v));
    b = neg(d.D3F(z, y, t, x, u, v));
    if (t == x) { s = cadd(s, d.D2(z, y, u, v)); b = cadd(b, d.D2F(z, y, u, v)); }
    b = cadd(b, cscale(s, e[u] + e[t] - e[v]));
  } else if (X == 'B') {  // solve_B_subspace.F90:317-345, 404-447
    const int x = pairs[2 * i], y = pairs[2 * i + 1], t = pairs[2 * j], u = pairs[2 * j + 1];
    s = d.D2(t, x, u, y);
    b = d.D2F(t, x, u, y);
    if (t == x) { s = csub(s, d.D1(u, y)); b = csub(b, d.D1F(u, y)); }
    if (u == y) { s = csub(s, d.D1(t, x)); b = csub(b, d.D1F(t, x)); }
    if (t == y) { s = csub(s, d.D1(u, x)); b = csub(b, d.D1F(u, x)); }
    if (t == x && u == y) { s.x += 1.0; b.x += e0; }
    b = cadd(b, cscale(s, e[t] + e[u]));
  } else if (X == 'C') {  // solve_C_subspace.F90:308-330, 384-414
    const int x = (int)(i / (na * na)), y = (int)((i / na) % na), z = (int)(i % na);
    const int t = (int)(j / (na * na)), u = (int)((j / na) % na), v = (int)(j % na);
    s = d.D3(z, y, x, t, u, v);
    b = cadd(d.D3F(z, y, x, t, u, v), cscale(s, e[u] - e[v] - e[t]));
  } else if (X == 'D') {  // solve_D_subspace.F90:326-337, 387-406
    const int x = (int)(i / na), y = (int)(i % na), t = (int)(j / na), u = (int)(j % na);
    s = d.D2(y, x, t, u);
    b = cadd(d.D2F(y, x, t, u), cscale(s, e[t] - e[u]));
  } else if (X == 'E') {  // solve_E_subspace.F90:303-318, 362-392
    const int u = (int)i, t = (int)j;
    s = neg(d.D1(t, u));
    if (u == t) s.x += 1.0;
    b = neg(d.D1F(t, u));
    if (u == t) b.x += e0;
    b = cadd(b, cscale(s, e[t]));
  } else if (X == 'F') {  // solve_F_subspace.F90:325-340, 395-420
    const int v = pairs[2 * i], x = pairs[2 * i + 1], t = pairs[2 * j], u = pairs[2 * j + 1];
    s = d.D2(v, t, x, u);
    b = d.D2F(v, t, x, u);
    if (x == t) { s = csub(s, d.D1(v, u)); b = csub(b, d.D1F(v, u)); }
    b = cadd(b, cscale(s, -e[u] - e[t]));
  } else {  // 'G', solve_G_subspace.F90:313-325, 380-395
    const int ti = (int)i, tj = (int)j;
    s = d.D1(ti, tj);
    b = csub(d.D1F(ti, tj), cscale(s, e[ti]));
  }
}

// map (optional, increasing): row/column k of the built matrices is tuple map[k] of the formula
template <char X>
__global__ void k_build_SB(DV d, const double* eps_act, const double* e0p, int64_t n, const int* pairs,
                           const int* map, cplx* S, cplx* B) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  const double e0 = *e0p;
  const int64_t fi = map ? map[i] : i, fj = map ? map[j] : j;
  cplx s, b;
  if (i <= j) {
    sb_formula<X>(d, eps_act, e0, pairs, fi, fj, s, b);
    if (i == j) { s = cconj(s); b = cconj(b); }   // sc(j,i) = DCONJG(sc(i,j)) with i == j
  } else {
    sb_formula<X>(d, eps_act, e0, pairs, fj, fi, s, b);
    s = cconj(s);
    b = cconj(b);
  }
  S[i + n * j] = s;
  B[i + n * j] = b;
}

// effh(p,i) = h'(p,i) + sum_k [(pi|kk) - (pk|ki)]      (solve_A_subspace.F90:531-612), A2 = (a,i|i,i)
__global__ void k_effh_A(int ni, int na, const cplx* hp, int ldh, const cplx* A2, cplx* effh) {
  const int p = threadIdx.x, i = blockIdx.x;
  if (p >= na) return;
  const int64_t Na = na, Ni = ni;
  cplx v = hp[(ni + p) + (int64_t)ldh * i];
  for (int k = 0; k < ni; ++k) {
    v = cadd(v, A2[p + Na * (i + Ni * (k + Ni * k))]);
    v = csub(v, A2[p + Na * (k + Ni * (k + Ni * i))]);
  }
  effh[p + na * i] = v;
}

// effh(a,p) = h'(a,p) + sum_k[(ap|kk) - (ak|kp)] - sum_w (aw|wp)   (solve_C_subspace.F90:530-608)
__global__ void k_effh_C(int ni, int na, int ns, const cplx* hp, int ldh, const cplx* C1, const cplx* C2,
                         const cplx* C3, cplx* effh) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x, p = blockIdx.y;
  if (a >= ns) return;
  const int64_t S = ns, Na = na, Ni = ni;
  cplx v = hp[(ni + na + a) + (int64_t)ldh * (ni + p)];
  for (int k = 0; k < ni; ++k) {
    v = cadd(v, C2[a + S * (p + Na * (k + Ni * k))]);
    v = csub(v, C3[a + S * (k + Ni * (k + Ni * p))]);
  }
  for (int w = 0; w < na; ++w) v = csub(v, C1[a + S * (w + Na * (w + Na * p))]);
  effh[a + ns * p] = v;
}

// effh(a,i) = h'(a,i) + sum_k [(ai|kk) - (ak|ki)]       (solve_D_subspace.F90:455-597), D3b = (s,i|i,i)
__global__ void k_effh_D(int ni, int na, int ns, const cplx* hp, int ldh, const cplx* D3b, cplx* effh) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (a >= ns) return;
  const int64_t S = ns, Ni = ni;
  cplx v = hp[(ni + na + a) + (int64_t)ldh * i];
  for (int k = 0; k < ni; ++k) {
    v = cadd(v, D3b[a + S * (i + Ni * (k + Ni * k))]);
    v = csub(v, D3b[a + S * (k + Ni * (k + Ni * i))]);
  }
  effh[a + ns * i] = v;
}

// Dm[(v + na u + na2 t) + na3 (p + na q + na2 r)] = D3(v,u,p,t,q,r)
__global__ void k_d3_perm_A(DV d, cplx* Dm) {
  const int na = d.na;
  const int64_t na3 = (int64_t)na * na * na;
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, col = blockIdx.y;
  if (row >= na3) return;
  const int v = (int)(row % na), u = (int)((row / na) % na), t = (int)(row / (na * na));
  const int p = (int)(col % na), q = (int)((col / na) % na), r = (int)(col / (na * na));
  Dm[row + na3 * col] = d.D3(v, u, p, t, q, r);
}
// Dm[(v + na u + na2 t) + na3 (j + na k + na2 l)] = D3(v,u,t,j,k,l)
__global__ void k_d3_perm_C(DV d, cplx* Dm) {
  const int na = d.na;
  const int64_t na3 = (int64_t)na * na * na;
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, col = blockIdx.y;
  if (row >= na3) return;
  const int v = (int)(row % na), u = (int)((row / na) % na), t = (int)(row / (na * na));
  const int j = (int)(col % na), k = (int)((col / na) % na), l = (int)(col / (na * na));
  Dm[row + na3 * col] = d.D3(v, u, t, j, k, l);
}
// The 3-RDM part of V_A and V_C without the 3-RDM.  With Y1[:, q + na r] = E_qr|0> over the
// full string space, <E_vu E_ts E_qr> = Y1[:, u + na v]^H E_ts Y1[:, q + na r], so the sum
// over (q, r) against an integral block is one product Zt = M Y1^T, the sum over the
// remaining index s goes through the excitation table (this kernel), and the contraction
// with <0|E_vu one more product (v3_permute_add).
//   SUMANN (V_C, sum_j <I|E_tj|Z_j>): the annihilation index s is summed, the image of I
//     under E_st is read;
//   !SUMANN (V_A, sum_p <I|E_pt|Z_p>): the creation index s is summed, E_ts's image.
// Wt[(x + nx t) + nx na I] = sum_s sign * Zt[x xs + s ss + ld J], J = that image of I.
template <bool SUMANN>
__global__ void k_v3_esum(int na, int64_t nf, int nx, int64_t xs, int64_t ss, int64_t ld, const int* __restrict__ tab,
                          const cplx* __restrict__ Zt, cplx* __restrict__ Wt) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)nx * nf) return;
  const int x = (int)(g % nx), t = blockIdx.y;
  const int64_t I = g / nx;
  cplx acc = mk(0.0, 0.0);
  for (int sidx = 0; sidx < na; ++sidx) {
    const int e = tab[(int64_t)(SUMANN ? sidx + na * t : t + na * sidx) * nf + I];
    if (e < 0) continue;
    const cplx v = Zt[(int64_t)x * xs + (int64_t)sidx * ss + ld * (int64_t)(e >> 1)];
    acc = (e & 1) ? csub(acc, v) : cadd(acc, v);
  }
  Wt[(int64_t)x + (int64_t)nx * t + (int64_t)nx * na * I] = acc;
}

// V[(v + na u + na2 t) + na3 x] = scale * Out[(u + na v) + na2 (x + nx t)]
__global__ void k_v3_permute(int na, int nx, double scale, const cplx* __restrict__ Out, cplx* __restrict__ V) {
  const int64_t na3 = (int64_t)na * na * na;
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, x = blockIdx.y;
  if (row >= na3) return;
  const int v = (int)(row % na), u = (int)((row / na) % na), t = (int)(row / (na * na));
  const cplx o = Out[(int64_t)(u + na * v) + (int64_t)na * na * (x + (int64_t)nx * t)];
  V[row + na3 * x] = mk(scale * o.x, scale * o.y);
}

// Am[(p + na q + na2 r) + na3 i] = A1[p + na (i + ni (q + na r))],  A1 = (a,i|a,a)
__global__ void k_a1_perm(int na, int ni, const cplx* A1, cplx* Am) {
  const int64_t na3 = (int64_t)na * na * na;
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (row >= na3) return;
  const int p = (int)(row % na), q = (int)((row / na) % na), r = (int)(row / (na * na));
  Am[row + na3 * i] = A1[p + (int64_t)na * (i + (int64_t)ni * (q + (int64_t)na * r))];
}

// V_A(tuv, i) += sum_pq <E_vu E_pq>(ti|pq) - sum_p <E_vu E_pt> effh(p,i) + <E_vu> effh(t,i)
__global__ void k_V_A_rest(int na, int ni, DV d, const cplx* A1, const cplx* effh, cplx* V) {
  const int64_t na3 = (int64_t)na * na * na;
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (row >= na3) return;
  const int v = (int)(row % na), u = (int)((row / na) % na), t = (int)(row / (na * na));
  const int64_t Na = na, Ni = ni;
  cplx acc = V[row + na3 * i];
  for (int q = 0; q < na; ++q)
    for (int p = 0; p < na; ++p) cfma(acc, d.D2(v, u, p, q), A1[t + Na * (i + Ni * (p + Na * q))]);
  for (int p = 0; p < na; ++p) acc = csub(acc, cmul(d.D2(v, u, p, t), effh[p + na * i]));
  cfma(acc, d.D1(v, u), effh[t + na * i]);
  V[row + na3 * i] = acc;
}

// V_C(tuv, a) += sum_p effh(a,p) <E_vu E_tp>, skipping |effh| < 1e-10 (solve_C_subspace.F90:629)
__global__ void k_V_C_rest(int na, int ns, DV d, const cplx* effh, cplx* V) {
  const int64_t na3 = (int64_t)na * na * na;
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, a = blockIdx.y;
  if (row >= na3) return;
  const int v = (int)(row % na), u = (int)((row / na) % na), t = (int)(row / (na * na));
  cplx acc = V[row + na3 * a];
  for (int p = 0; p < na; ++p) {
    const cplx ef = effh[a + (int64_t)ns * p];
    if (sqrt(cabs2(ef)) < 1.0e-10) continue;
    cfma(acc, ef, d.D2(v, u, t, p));
  }
  V[row + na3 * a] = acc;
}

// V_B(tu; i>j)  (solve_B_subspace.F90:530-585), Bb = (a,i|a,i)
__global__ void k_V_B(int na, int ni, DV d, const cplx* Bb, const int* tp, const int* ip, int64_t ntp, cplx* V) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, m = blockIdx.y;
  if (k >= ntp) return;
  const int t = tp[2 * k], u = tp[2 * k + 1], i = ip[2 * m], j = ip[2 * m + 1];
  const int64_t Na = na, Ni = ni;
  auto Bv = [&](int p, int ii, int q, int jj) { return Bb[p + Na * (ii + Ni * (q + Na * jj))]; };
  cplx acc = csub(Bv(t, i, u, j), Bv(u, i, t, j));
  for (int q = 0; q < na; ++q)
    for (int p = 0; p < na; ++p) cfma(acc, Bv(p, i, q, j), d.D2(p, t, q, u));
  for (int p = 0; p < na; ++p) {
    cfma(acc, Bv(u, i, p, j), d.D1(p, t));
    acc = csub(acc, cmul(Bv(t, i, p, j), d.D1(p, u)));
    acc = csub(acc, cmul(Bv(p, i, u, j), d.D1(p, t)));
  }
  V[k + ntp * m] = acc;
}

// V_D(tu; a,i) (solve_D_subspace.F90:455-597), D1b = (s,i|a,a), D2b = (s,a|a,i)
__global__ void k_V_D(int na, int ni, int ns, DV d, const cplx* D1b, const cplx* D2b, const cplx* effh, cplx* V) {
  const int64_t tu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, ai = blockIdx.y;
  const int64_t na2 = (int64_t)na * na;
  if (tu >= na2) return;
  const int t = (int)(tu / na), u = (int)(tu % na), a = (int)(ai / ni), i = (int)(ai % ni);
  const int64_t S = ns, Na = na, Ni = ni;
  cplx acc = cmul(effh[a + S * i], d.D1(u, t));
  for (int q = 0; q < na; ++q)
    for (int p = 0; p < na; ++p) {
      const cplx g = csub(D1b[a + S * (i + Ni * (p + Na * q))], D2b[a + S * (q + Na * (p + Na * i))]);
      cfma(acc, d.D2(u, t, p, q), g);
    }
  V[tu + na2 * ai] = acc;
}

// V_E(t; a,I>J) (solve_E_subspace.F90:436-466), Eb = (s,i|a,i)
__global__ void k_V_E(int na, int ni, int ns, DV d, const cplx* Eb, const int* ip, cplx* V) {
  const int t = threadIdx.x;
  const int64_t pa = blockIdx.x;   // pair * ns + a
  if (t >= na) return;
  const int64_t S = ns, Ni = ni, Na = na;
  const int a = (int)(pa % ns), m = (int)(pa / ns);
  const int I = ip[2 * m], J = ip[2 * m + 1];
  auto Ev = [&](int aa, int ii, int p, int jj) { return Eb[aa + S * (ii + Ni * (p + Na * jj))]; };
  cplx acc = csub(Ev(a, J, t, I), Ev(a, I, t, J));
  for (int p = 0; p < na; ++p) cfma(acc, d.D1(t, p), csub(Ev(a, I, p, J), Ev(a, J, p, I)));
  V[t + Na * pa] = acc;
}

// V_F(tu; a>b) (solve_F_subspace.F90:505-535), Fb = (s,a|s,a): sum_pq D2(t,p,u,q) F(a,p,b,q)
// - sum_p D1(t,p) F(a,u,b,p), as a matrix product: with the pair index k = (t > u),
// Am[k, (p,q)] = D2(t,p,u,q) - [p = u] D1(t,q) and Fp[(p,q), (a,b)] = F(a,p,b,q),
// C = Am Fp holds V_F(k; a, b) for every a, b; the a > b columns are gathered into V.
__global__ void k_vf_amat(int na, DV d, const int* tp, int64_t ntp, cplx* Am) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t na2 = (int64_t)na * na;
  if (i >= ntp * na2) return;
  const int64_t k = i % ntp, pq = i / ntp;
  const int p = (int)(pq % na), q = (int)(pq / na), t = tp[2 * k], u = tp[2 * k + 1];
  cplx v = d.D2(t, p, u, q);
  if (p == u) v = csub(v, d.D1(t, q));
  Am[i] = v;
}

// Fp[(p + na q) + na^2 (a + ns b)] = Fb[a + ns (p + na (b + ns q))]: for each b a transpose of
// (a) x (p,q) through a 32 x 32 LDS tile, both sides coalesced
__global__ __launch_bounds__(256) void k_vf_perm(int na, int ns, const cplx* __restrict__ Fb, cplx* __restrict__ Fp) {
  __shared__ cplx tile[32][33];
  const int b = blockIdx.z, a0 = blockIdx.x * 32, pq0 = blockIdx.y * 32;
  const int na2 = na * na, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
  const int64_t S = ns, Na = na;
  for (int r = ty; r < 32; r += 8) {   // read: a = a0 + tx contiguous, pq = pq0 + r
    const int a = a0 + tx, pq = pq0 + r;
    if (a < ns && pq < na2) {
      const int p = pq % na, q = pq / na;
      tile[r][tx] = Fb[a + S * (p + Na * (b + S * q))];
    }
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {   // write: pq = pq0 + tx contiguous, a = a0 + r
    const int a = a0 + r, pq = pq0 + tx;
    if (a < ns && pq < na2) Fp[pq + (int64_t)na2 * (a + S * b)] = tile[tx][r];
  }
}

__global__ void k_vf_gather(const cplx* __restrict__ C, int ns, const int* sp, int64_t ntp, int64_t nsp,
                            cplx* __restrict__ V) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntp * nsp) return;
  const int64_t k = i % ntp, m = i / ntp;
  const int64_t ab = sp[2 * m] + (int64_t)ns * sp[2 * m + 1];
  V[i] = C[k + ntp * ab];
}

// V_G(t; a>b, i) (solve_G_subspace.F90:445-462), Gb = (s,i|s,a)
__global__ void k_V_G(int na, int ni, int ns, DV d, const cplx* Gb, const int* sp, int64_t nsp, cplx* V) {
  const int t = threadIdx.x % 32 < na ? threadIdx.x % 32 : -1;
  const int64_t pi = (int64_t)blockIdx.x * (blockDim.x / 32) + threadIdx.x / 32;  // pair*ni + i
  if (t < 0 || pi >= nsp * ni) return;
  const int64_t S = ns, Ni = ni;
  const int m = (int)(pi / ni), i = (int)(pi % ni), a = sp[2 * m], b = sp[2 * m + 1];
  auto Gv = [&](int aa, int ii, int bb, int p) { return Gb[aa + S * (ii + Ni * (bb + S * p))]; };
  cplx acc = mk(0.0, 0.0);
  for (int p = 0; p < na; ++p) cfma(acc, d.D1(t, p), csub(Gv(a, i, b, p), Gv(b, i, a, p)));
  V[t + (int64_t)na * pi] = acc;
}

__global__ void k_alphas(char X, int ni, int na, int ns, const double* eps, const double* e0p, double eshift,
                         const int* ip, const int* sp, int64_t npert, double* alpha) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npert) return;
  const double e0 = *e0p;
  const int nmoc = ni + na;
  double a = 0.0;
  switch (X) {
    case 'A': a = -eps[k]; break;                                              // :249
    case 'B': a = -eps[ip[2 * k]] - eps[ip[2 * k + 1]]; break;                 // :251
    case 'C': a = eps[nmoc + k]; break;                                        // :247
    case 'D': a = eps[nmoc + k / ni] - eps[k % ni]; break;                     // :262
    case 'E': { const int64_t m = k / ns, s = k % ns;                          // :238
      a = eps[nmoc + s] - eps[ip[2 * m]] - eps[ip[2 * m + 1]]; } break;
    case 'F': a = eps[nmoc + sp[2 * k]] + eps[nmoc + sp[2 * k + 1]]; break;    // :262
    case 'G': { const int64_t m = k / ni, i = k % ni;                          // :252
      a = -eps[i] + eps[nmoc + sp[2 * m]] + eps[nmoc + sp[2 * m + 1]]; } break;
  }
  alpha[k] = a - e0 + eshift;
}

constexpr int RB = 256;
__device__ inline void block_sum2(double& a, double& b, double* sh) {
  const int t = threadIdx.x;
  sh[t] = a;
  sh[RB + t] = b;
  __syncthreads();
  for (int o = RB / 2; o > 0; o >>= 1) {
    if (t < o) { sh[t] += sh[t + o]; sh[RB + t] += sh[RB + t + o]; }
    __syncthreads();
  }
  a = sh[0];
  b = sh[RB];
}

__global__ __launch_bounds__(RB) void k_e2_partial(int64_t dimm, int64_t npert, const cplx* X, const double* wb,
                                                   const double* alpha, double* partial) {
  __shared__ double sh[2 * RB];
  double e = 0.0, s = 0.0;
  const int64_t total = dimm * npert;
  for (int64_t q = (int64_t)blockIdx.x * RB + threadIdx.x; q < total; q += (int64_t)gridDim.x * RB) {
    const int64_t j = q % dimm, k = q / dimm;
    const double den = alpha[k] + wb[j];
    const double a2 = cabs2(X[q]);
    e -= a2 / den;
    s += a2 / (den * den);
  }
  block_sum2(e, s, sh);
  if (threadIdx.x == 0) { partial[2 * blockIdx.x] = e; partial[2 * blockIdx.x + 1] = s; }
}

__global__ __launch_bounds__(RB) void k_e2_H_partial(int ni, int ns, const cplx* Hb, const double* eps, double eshift,
                                                     const int* ip, const int* sp, int64_t nsp, int64_t nip,
                                                     double* partial) {
  __shared__ double sh[2 * RB];
  double e = 0.0, s = 0.0;
  const int64_t S = ns, Ni = ni, total = nsp * nip;
  for (int64_t q = (int64_t)blockIdx.x * RB + threadIdx.x; q < total; q += (int64_t)gridDim.x * RB) {
    const int64_t m = q % nsp, n = q / nsp;
    const int a = sp[2 * m], b = sp[2 * m + 1], i = ip[2 * n], j = ip[2 * n + 1];
    const cplx v = csub(Hb[a + S * (i + Ni * (b + S * j))], Hb[a + S * (j + Ni * (b + S * i))]);
    const double den = eps[a] + eps[b] - eps[ns + i] - eps[ns + j] + eshift;
    const double a2 = cabs2(v);
    e -= a2 / den;
    s += a2 / (den * den);
  }
  block_sum2(e, s, sh);
  if (threadIdx.x == 0) { partial[2 * blockIdx.x] = e; partial[2 * blockIdx.x + 1] = s; }
}

// E2 of one perturber k through the compressed resolvent (one workgroup per k).  With
// G = [Gd | Gk] unitary (Gd: the r dropped directions), D = diag(f + alpha_k) and
// z = Gk y_k, the retained-space resolvent gives u = (Gk^H D Gk)^-1 y_k = Gk^H x with
//   x = D^-1 (z - Gd c),  (Gd^H D^-1 Gd) c = Gd^H D^-1 z,
// so e2_k = -Re(z^H x) = -y^H (B~ + alpha_k)^-1 y and sumc2_k = |x|^2.  The r x r system is
// solved in LDS by Gaussian elimination with partial pivoting; a zero pivot yields NaN,
// which the host detects.
constexpr int SCHUR_RMAX = 32;
__global__ __launch_bounds__(RB) void k_e2_schur(int64_t m, int r, const cplx* __restrict__ Gd,
                                                 const cplx* __restrict__ Z, const double* __restrict__ f,
                                                 const double* __restrict__ alpha, double* partial) {
  __shared__ cplx W[SCHUR_RMAX][SCHUR_RMAX + 1];   // column r holds the right-hand side
  __shared__ cplx lf[SCHUR_RMAX], cs[SCHUR_RMAX];
  __shared__ double sh[2 * RB];
  __shared__ int piv;
  const int t = threadIdx.x;
  const int64_t k = blockIdx.x;
  const double al = alpha[k];
  const cplx* z = Z + m * k;
  // W(a,b) = sum_row conj(Gd(row,a)) Gd(row,b)/D(row), W(a,r) = sum_row conj(Gd(row,a)) z(row)/D(row)
  for (int o = t; o < r * (r + 1); o += RB) {
    const int a = o % r, b = o / r;
    const cplx* gb = b < r ? Gd + m * b : z;
    const cplx* ga = Gd + m * a;
    cplx acc = mk(0.0, 0.0);
    for (int64_t row = 0; row < m; ++row) cfma(acc, cconj(ga[row]), cscale(gb[row], 1.0 / (f[row] + al)));
    W[a][b] = acc;
  }
  __syncthreads();
  for (int j = 0; j < r; ++j) {
    if (t == 0) {
      int p = j;
      double best = cabs2(W[j][j]);
      for (int i = j + 1; i < r; ++i)
        if (cabs2(W[i][j]) > best) { best = cabs2(W[i][j]); p = i; }
      piv = p;
    }
    __syncthreads();
    if (piv != j)
      for (int b = t; b <= r; b += RB) { const cplx tmp = W[j][b]; W[j][b] = W[piv][b]; W[piv][b] = tmp; }
    __syncthreads();
    const cplx d = W[j][j];
    const double dd = cabs2(d);
    for (int i = j + 1 + t; i < r; i += RB)
      lf[i] = dd > 0.0 ? cscale(cmul(W[i][j], cconj(d)), 1.0 / dd) : mk(NAN, NAN);
    __syncthreads();
    for (int o = t; o < (r - j - 1) * (r - j); o += RB) {
      const int i = j + 1 + o % (r - j - 1), b = j + 1 + o / (r - j - 1);
      W[i][b] = csub(W[i][b], cmul(lf[i], W[j][b]));
    }
    __syncthreads();
  }
  if (t == 0) {
    for (int i = r - 1; i >= 0; --i) {
      cplx acc = W[i][r];
      for (int b = i + 1; b < r; ++b) acc = csub(acc, cmul(W[i][b], cs[b]));
      const double dd = cabs2(W[i][i]);
      cs[i] = dd > 0.0 ? cscale(cmul(acc, cconj(W[i][i])), 1.0 / dd) : mk(NAN, NAN);
    }
  }
  __syncthreads();
  double e = 0.0, s = 0.0;
  for (int64_t row = t; row < m; row += RB) {
    cplx acc = z[row];
    for (int a = 0; a < r; ++a) acc = csub(acc, cmul(Gd[row + m * a], cs[a]));
    const cplx x = cscale(acc, 1.0 / (f[row] + al));
    const cplx zx = cmul(cconj(z[row]), x);
    e -= zx.x;
    s += cabs2(x);
  }
  block_sum2(e, s, sh);
  if (t == 0) { partial[2 * k] = e; partial[2 * k + 1] = s; }
}

__global__ __launch_bounds__(RB) void k_final2(int64_t np, const double* partial, double* out2) {
  __shared__ double sh[2 * RB];
  double e = 0.0, s = 0.0;
  for (int64_t q = threadIdx.x; q < np; q += RB) { e += partial[2 * q]; s += partial[2 * q + 1]; }
  block_sum2(e, s, sh);
  if (threadIdx.x == 0) { out2[0] = e; out2[1] = s; }
}

__global__ void k_gather_rows(int64_t nr, int64_t ncol, const int* map, const cplx* in, int64_t ldin, cplx* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (i < nr) out[i + nr * c] = in[map[i] + ldin * c];
}

}  // namespace

int64_t e2_partials_needed(int64_t nwork) {
  int64_t nb = (nwork + RB - 1) / RB;
  if (nb > 2048) nb = 2048;
  if (nb < 1) nb = 1;
  return nb;
}

void build_SB(hipStream_t s, char X, const Dens& d, const double* eps_act, const double* e0_dev, int64_t n,
              const int* pairs, cplx* S, cplx* B, const int* map) {
  DC2_CHECK(n < 65536, 1, "build_SB: dimension too large");
  dim3 g(cdiv(n, 128), (unsigned)n);
  DV v = dv(d);
  switch (X) {
    case 'A': hipLaunchKernelGGL(k_build_SB<'A'>, g, dim3(128), 0, s, v, eps_act, e0_dev, n, pairs, map, S, B); break;
    case 'B': hipLaunchKernelGGL(k_build_SB<'B'>, g, dim3(128), 0, s, v, eps_act, e0_dev, n, pairs, map, S, B); break;
    case 'C': hipLaunchKernelGGL(k_build_SB<'C'>, g, dim3(128), 0, s, v, eps_act, e0_dev, n, pairs, map, S, B); break;
    case 'D': hipLaunchKernelGGL(k_build_SB<'D'>, g, dim3(128), 0, s, v, eps_act, e0_dev, n, pairs, map, S, B); break;
    case 'E': hipLaunchKernelGGL(k_build_SB<'E'>, g, dim3(128), 0, s, v, eps_act, e0_dev, n, pairs, map, S, B); break;
    case 'F': hipLaunchKernelGGL(k_build_SB<'F'>, g, dim3(128), 0, s, v, eps_act, e0_dev, n, pairs, map, S, B); break;
    case 'G': hipLaunchKernelGGL(k_build_SB<'G'>, g, dim3(128), 0, s, v, eps_act, e0_dev, n, pairs, map, S, B); break;
    default: throw Error(1, "build_SB: bad subspace");
  }
  DC2_HIP(hipGetLastError());
}

void gather_rows(hipStream_t s, int64_t nr, int64_t ncol, const int* map, const cplx* in, int64_t ldin, cplx* out) {
  if (nr <= 0 || ncol <= 0) return;
  DC2_CHECK(ncol < 65536, 1, "gather_rows: too many columns");
  hipLaunchKernelGGL(k_gather_rows, dim3(cdiv(nr, 256), (unsigned)ncol), dim3(256), 0, s, nr, ncol, map, in, ldin,
                     out);
  DC2_HIP(hipGetLastError());
}

void effh_A(hipStream_t s, int ni, int na, const cplx* hp, int ldh, const cplx* A2, cplx* effh) {
  hipLaunchKernelGGL(k_effh_A, dim3(ni), dim3(64), 0, s, ni, na, hp, ldh, A2, effh);
  DC2_HIP(hipGetLastError());
}
void effh_C(hipStream_t s, int ni, int na, int ns, const cplx* hp, int ldh, const cplx* C1, const cplx* C2,
            const cplx* C3, cplx* effh) {
  hipLaunchKernelGGL(k_effh_C, dim3(cdiv(ns, 64), na), dim3(64), 0, s, ni, na, ns, hp, ldh, C1, C2, C3, effh);
  DC2_HIP(hipGetLastError());
}
void effh_D(hipStream_t s, int ni, int na, int ns, const cplx* hp, int ldh, const cplx* D3b, cplx* effh) {
  hipLaunchKernelGGL(k_effh_D, dim3(cdiv(ns, 64), ni), dim3(64), 0, s, ni, na, ns, hp, ldh, D3b, effh);
  DC2_HIP(hipGetLastError());
}
void d3_perm_A(hipStream_t s, int na, const cplx* G3, cplx* Dm) {
  Dens d{na, nullptr, nullptr, nullptr, nullptr, G3, nullptr};
  const int64_t na3 = (int64_t)na * na * na;
  hipLaunchKernelGGL(k_d3_perm_A, dim3(cdiv(na3, 256), (unsigned)na3), dim3(256), 0, s, dv(d), Dm);
  DC2_HIP(hipGetLastError());
}
void d3_perm_C(hipStream_t s, int na, const cplx* G3, cplx* Dm) {
  Dens d{na, nullptr, nullptr, nullptr, nullptr, G3, nullptr};
  const int64_t na3 = (int64_t)na * na * na;
  hipLaunchKernelGGL(k_d3_perm_C, dim3(cdiv(na3, 256), (unsigned)na3), dim3(256), 0, s, dv(d), Dm);
  DC2_HIP(hipGetLastError());
}
void v3_esum(hipStream_t s, bool sum_annihilation, int na, int64_t nf, int nx, int64_t xs, int64_t ss, int64_t ld,
             const int* tab, const cplx* Zt, cplx* Wt) {
  if (nx <= 0 || nf <= 0) return;
  const dim3 grid((unsigned)cdiv((int64_t)nx * nf, 256), (unsigned)na);
  if (sum_annihilation)
    hipLaunchKernelGGL(k_v3_esum<true>, grid, dim3(256), 0, s, na, nf, nx, xs, ss, ld, tab, Zt, Wt);
  else
    hipLaunchKernelGGL(k_v3_esum<false>, grid, dim3(256), 0, s, na, nf, nx, xs, ss, ld, tab, Zt, Wt);
  DC2_HIP(hipGetLastError());
}
void v3_permute(hipStream_t s, int na, int nx, double scale, const cplx* Out, cplx* V) {
  const int64_t na3 = (int64_t)na * na * na;
  if (nx <= 0) return;
  DC2_CHECK(nx < 65536, 1, "v3_permute: too many columns");
  hipLaunchKernelGGL(k_v3_permute, dim3(cdiv(na3, 256), (unsigned)nx), dim3(256), 0, s, na, nx, scale, Out, V);
  DC2_HIP(hipGetLastError());
}
void a1_perm(hipStream_t s, int na, int ni, const cplx* A1, cplx* Am) {
  const int64_t na3 = (int64_t)na * na * na;
  hipLaunchKernelGGL(k_a1_perm, dim3(cdiv(na3, 256), ni), dim3(256), 0, s, na, ni, A1, Am);
  DC2_HIP(hipGetLastError());
}
void V_A_rest(hipStream_t s, int na, int ni, const Dens& d, const cplx* A1, const cplx* effh, cplx* V) {
  const int64_t na3 = (int64_t)na * na * na;
  hipLaunchKernelGGL(k_V_A_rest, dim3(cdiv(na3, 128), ni), dim3(128), 0, s, na, ni, dv(d), A1, effh, V);
  DC2_HIP(hipGetLastError());
}
void V_C_rest(hipStream_t s, int na, int ns, const Dens& d, const cplx* effh, cplx* V) {
  const int64_t na3 = (int64_t)na * na * na;
  hipLaunchKernelGGL(k_V_C_rest, dim3(cdiv(na3, 128), ns), dim3(128), 0, s, na, ns, dv(d), effh, V);
  DC2_HIP(hipGetLastError());
}
void V_B(hipStream_t s, int na, int ni, const Dens& d, const cplx* Bb, const int* tpairs, const int* ipairs,
         cplx* V) {
  const int64_t ntp = (int64_t)na * (na - 1) / 2, nip = (int64_t)ni * (ni - 1) / 2;
  if (ntp == 0 || nip == 0) return;
  hipLaunchKernelGGL(k_V_B, dim3(cdiv(ntp, 64), (unsigned)nip), dim3(64), 0, s, na, ni, dv(d), Bb, tpairs, ipairs,
                     ntp, V);
  DC2_HIP(hipGetLastError());
}
void V_D(hipStream_t s, int na, int ni, int ns, const Dens& d, const cplx* D1b, const cplx* D2b, const cplx* effh,
         cplx* V) {
  const int64_t na2 = (int64_t)na * na;
  hipLaunchKernelGGL(k_V_D, dim3(cdiv(na2, 64), (unsigned)((int64_t)ns * ni)), dim3(64), 0, s, na, ni, ns, dv(d),
                     D1b, D2b, effh, V);
  DC2_HIP(hipGetLastError());
}
void V_E(hipStream_t s, int na, int ni, int ns, const Dens& d, const cplx* Eb, const int* ipairs, cplx* V) {
  const int64_t nip = (int64_t)ni * (ni - 1) / 2;
  if (nip == 0) return;
  DC2_CHECK(na <= 64, 1, "V_E: nact <= 64");
  hipLaunchKernelGGL(k_V_E, dim3((unsigned)(nip * ns)), dim3(64), 0, s, na, ni, ns, dv(d), Eb, ipairs, V);
  DC2_HIP(hipGetLastError());
}
int64_t V_F_scratch(int na, int ns) {
  const int64_t ntp = (int64_t)na * (na - 1) / 2, na2 = (int64_t)na * na, ns2 = (int64_t)ns * ns;
  return na2 * ns2 + ntp * na2 + ntp * ns2;
}

void V_F(hipStream_t s, int na, int ns, const Dens& d, const cplx* Fb, const int* tpairs, const int* spairs,
         cplx* V, cplx* scratch) {
  const int64_t ntp = (int64_t)na * (na - 1) / 2, nsp = (int64_t)ns * (ns - 1) / 2;
  if (ntp == 0 || nsp == 0) return;
  const int64_t na2 = (int64_t)na * na, ns2 = (int64_t)ns * ns;
  DC2_CHECK(ns < 65536 && cdiv(na2, 32) < 65536, 1, "V_F: too many orbitals for the grid");
  cplx* Fp = scratch;
  cplx* Am = Fp + na2 * ns2;
  cplx* C = Am + ntp * na2;
  hipLaunchKernelGGL(k_vf_amat, dim3(cdiv(ntp * na2, 256)), dim3(256), 0, s, na, dv(d), tpairs, ntp, Am);
  hipLaunchKernelGGL(k_vf_perm, dim3(cdiv(ns, 32), cdiv(na2, 32), (unsigned)ns), dim3(256), 0, s, na, ns, Fb, Fp);
  DC2_HIP(hipGetLastError());
  zgemm(s, 'N', 'N', ntp, ns2, na2, mk(1.0, 0.0), Am, ntp, 0, Fp, na2, 0, mk(0.0, 0.0), C, ntp, 0, 1);
  hipLaunchKernelGGL(k_vf_gather, dim3(cdiv(ntp * nsp, 256)), dim3(256), 0, s, C, ns, spairs, ntp, nsp, V);
  DC2_HIP(hipGetLastError());
}
void V_G(hipStream_t s, int na, int ni, int ns, const Dens& d, const cplx* Gb, const int* spairs, cplx* V) {
  const int64_t nsp = (int64_t)ns * (ns - 1) / 2;
  if (nsp == 0 || ni == 0) return;
  DC2_CHECK(na <= 32, 1, "V_G: nact <= 32");
  const int64_t items = nsp * ni;   // 8 items of 32 lanes per 256-thread block
  hipLaunchKernelGGL(k_V_G, dim3(cdiv(items, 8)), dim3(256), 0, s, na, ni, ns, dv(d), Gb, spairs, nsp, V);
  DC2_HIP(hipGetLastError());
}
void alphas(hipStream_t s, char X, int ni, int na, int ns, const double* eps, const double* e0_dev, double eshift,
            const int* ipairs, const int* spairs, int64_t npert, double* alpha) {
  if (npert <= 0) return;
  hipLaunchKernelGGL(k_alphas, dim3(cdiv(npert, 256)), dim3(256), 0, s, X, ni, na, ns, eps, e0_dev, eshift, ipairs,
                     spairs, npert, alpha);
  DC2_HIP(hipGetLastError());
}
void e2_reduce(hipStream_t s, int64_t dimm, int64_t npert, const cplx* X, const double* wb, const double* alpha,
               double* partial, int64_t npartial_max, double* out2) {
  int64_t nb = e2_partials_needed(dimm * npert);
  if (nb > npartial_max) nb = npartial_max;
  hipLaunchKernelGGL(k_e2_partial, dim3((unsigned)nb), dim3(RB), 0, s, dimm, npert, X, wb, alpha, partial);
  hipLaunchKernelGGL(k_final2, dim3(1), dim3(RB), 0, s, nb, partial, out2);
  DC2_HIP(hipGetLastError());
}
int e2_schur_max_dropped() { return SCHUR_RMAX; }
void e2_schur(hipStream_t s, int64_t m, int r, int64_t npert, const cplx* Gd, const cplx* Z, const double* f,
              const double* alpha, double* partial, double* out2) {
  DC2_CHECK(r >= 0 && r <= SCHUR_RMAX, 1, "e2_schur: too many dropped directions");
  DC2_CHECK(npert >= 1 && npert < (1ll << 31), 1, "e2_schur: perturber count out of range");
  hipLaunchKernelGGL(k_e2_schur, dim3((unsigned)npert), dim3(RB), 0, s, m, r, Gd, Z, f, alpha, partial);
  hipLaunchKernelGGL(k_final2, dim3(1), dim3(RB), 0, s, npert, partial, out2);
  DC2_HIP(hipGetLastError());
}
void e2_H(hipStream_t s, int ni, int ns, const cplx* Hb, const double* eps, double eshift, const int* ipairs,
          const int* spairs, double* partial, int64_t npartial_max, double* out2) {
  const int64_t nsp = (int64_t)ns * (ns - 1) / 2, nip = (int64_t)ni * (ni - 1) / 2;
  int64_t nb = e2_partials_needed(nsp * nip);
  if (nb > npartial_max) nb = npartial_max;
  // eps here is laid out [secondary..., inactive...] by the caller
  hipLaunchKernelGGL(k_e2_H_partial, dim3((unsigned)nb), dim3(RB), 0, s, ni, ns, Hb, eps, eshift, ipairs, spairs, nsp,
                     nip, partial);
  hipLaunchKernelGGL(k_final2, dim3(1), dim3(RB), 0, s, nb, partial, out2);
  DC2_HIP(hipGetLastError());
}

}  // namespace dc2
