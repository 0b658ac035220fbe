// CASCI-side kernels: determinant strings, E_pq tables and gathers, CI matrix,
// Fock matrix, eigenvector gauge, CI-vector rotation determinants.
//
// Strings are int64 bit masks over the active spinors (bit w-1 <-> active
// spinor w), exactly the integers search_cas_configuration.F90 stores in
// dict_cas_idx; all string work is integer arithmetic and bit-exact.
#include "dc2_kernels.hpp"

namespace dc2 {
namespace {

__device__ inline int64_t one_e_exct(int64_t det, int cre, int ann, int nact, int& phase) {
  // one_e_exct.F90:34-47, 1-based orbital labels
  const uint64_t ba = 1ull << (ann - 1), bc = 1ull << (cre - 1);
  const uint64_t d = (uint64_t)det;
  if (ann == cre && (d & ba)) { phase = 0; return det; }
  if ((d & ba) && !(d & bc)) {
    const uint64_t full = (nact >= 64) ? ~0ull : ((1ull << nact) - 1ull);
    const uint64_t ia = ann < nact ? (full & ~((1ull << ann) - 1ull)) : 0ull;
    const uint64_t ib = cre < nact ? (full & ~((1ull << cre) - 1ull)) : 0ull;
    phase = __popcll(d & ia) + __popcll((d - ba) & ib);
    return (int64_t)(d - ba + bc);
  }
  phase = 0;
  return 0;
}

__device__ inline bool ras_ok(uint64_t det, const DetParams& p) {
  // ras_det_check.F90:8-39
  if (p.ras1 != 0) {
    const int n = __popcll(det & ((1ull << p.ras1) - 1ull));
    if (!((p.ras1 - p.ras1_max_hole) <= n && n <= (p.ras1 - p.min_hole_ras1))) return false;
  }
  if (p.ras3 != 0) {
    const int sh = p.ras1 + p.ras2;
    const uint64_t r3 = (det >> sh) << sh;
    const int n = __popcll(det & r3);
    if (!(n <= p.ras3_max_elec && p.min_elec_ras3 <= n)) return false;
  }
  return true;
}

__device__ inline int det_totsym(uint64_t det, const DetParams& p) {
  // cas_configuration_totsym, search_cas_configuration.F90:112-149
  int ret = 1, ielec = 0;
  for (int i = 1; i <= p.nact; ++i) {
    if ((det >> (i - 1)) & 1ull) {
      ++ielec;
      const int jsym = p.irp_act[i - 1];
      if (ielec & 1) {
        ret = p.multb_ds[(jsym - 1) + p.nsym * (ret - 1)];
      } else {
        const int q = (jsym & 1) ? jsym + 1 : jsym - 1;
        ret = p.multb_d[(q - 1) + p.nsym * (ret - 1)];
      }
    }
  }
  if ((ielec & 1) == 0) ret += p.nsym;
  return ret;
}

__global__ void k_det_flags(DetParams p, int64_t N, int* ff, int* fc) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  const bool full = __popcll((uint64_t)x) == p.nelec;
  bool cas = full && ras_ok((uint64_t)x, p);
  if (cas && p.nsym != 1) cas = det_totsym((uint64_t)x, p) == p.totsym;
  ff[x] = full ? 1 : 0;
  fc[x] = cas ? 1 : 0;
}

// exclusive scans of both flag arrays in one workgroup (N <= 2^24)
__global__ __launch_bounds__(1024) void k_scan2(int64_t N, const int* ff, const int* fc, int* rf, int* rc,
                                                int* counts) {
  __shared__ int sf[1024], sc[1024];
  const int t = threadIdx.x;
  const int64_t chunk = (N + 1023) / 1024;
  const int64_t lo = t * chunk, hi = (lo + chunk < N) ? lo + chunk : N;
  int a = 0, b = 0;
  for (int64_t x = lo; x < hi; ++x) { a += ff[x]; b += fc[x]; }
  sf[t] = a;
  sc[t] = b;
  __syncthreads();
  if (t == 0) {
    int ra = 0, rb = 0;
    for (int k = 0; k < 1024; ++k) {
      const int va = sf[k], vb = sc[k];
      sf[k] = ra; sc[k] = rb;
      ra += va; rb += vb;
    }
    counts[0] = ra;
    counts[1] = rb;
  }
  __syncthreads();
  int ra = sf[t], rb = sc[t];
  for (int64_t x = lo; x < hi; ++x) {
    rf[x] = ff[x] ? ra++ : -1;
    rc[x] = fc[x] ? rb++ : -1;
  }
}

__global__ void k_det_scatter(int64_t N, const int* rf, const int* rc, int64_t* full_dets, int64_t* cas_dets,
                              int* cas2full) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  if (rf[x] >= 0) full_dets[rf[x]] = x;
  if (rc[x] >= 0) { cas_dets[rc[x]] = x; cas2full[rc[x]] = rf[x]; }
}

__global__ void k_exc_table(int na, int64_t nf, const int64_t* dets, const int* rank_full, int* tab) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nf) return;
  const int pq = blockIdx.y, p = pq % na, q = pq / na;
  int ph;
  const int64_t nw = one_e_exct(dets[j], p + 1, q + 1, na, ph);
  tab[(int64_t)pq * nf + j] = nw ? ((rank_full[nw] << 1) | (ph & 1)) : -1;
}

__global__ void k_apply_E(int na, int64_t nf, const int* tab, const cplx* __restrict__ X, cplx* __restrict__ Y) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nf) return;
  const int pq = blockIdx.y, p = pq % na, q = pq / na;
  const int64_t x = blockIdx.z;
  const int na2 = na * na;
  // <j|E_pq|i> = <i|E_qp|j>: gather from the E_qp image of j
  const int e = tab[(int64_t)(q + na * p) * nf + j];
  cplx v = mk(0.0, 0.0);
  if (e >= 0) {
    v = X[(int64_t)(e >> 1) + nf * x];
    if (e & 1) v = mk(-v.x, -v.y);
  }
  Y[j + nf * ((int64_t)pq + na2 * x)] = v;
}

__global__ void k_embed(int na, int64_t ncas, const int* cas2full, const cplx* c_cas, cplx* c_full) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ncas) c_full[cas2full[k]] = c_cas[k];
}

__global__ void k_fvec(int na, int64_t nf, const int64_t* dets, const double* eps_act, const cplx* c_full,
                       cplx* cF) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nf) return;
  double f = 0.0;
  const uint64_t d = (uint64_t)dets[j];
  for (int w = 0; w < na; ++w)
    if ((d >> w) & 1ull) f += eps_act[w];
  cF[j] = cscale(c_full[j], f);
}

__device__ inline cplx wave_sum(cplx v) {
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_down(v.x, o, 64);
    v.y += __shfl_down(v.y, o, 64);
  }
  return v;
}

// One wave per CI row (casmat.F90:37-216)
__global__ __launch_bounds__(64) void k_casmat(int ni, int na, int ne, int64_t ndet, const int64_t* dets,
                                               const int* rank_cas, const cplx* __restrict__ O, int nmoc,
                                               const cplx* __restrict__ h, int ldh, cplx* mat) {
  const int64_t i = blockIdx.x;
  const int lane = threadIdx.x;
  __shared__ int oc[32], vi[32], L[96];
  __shared__ short op0[512], op1[512], vp0[512], vp1[512];
  __shared__ int nop_s, nvp_s;
  const int64_t det = dets[i];
  const int nv = na - ne;
  if (lane == 0) {
    int no = 0, nvv = 0;
    for (int imo = 1; imo <= na; ++imo) {
      if ((det >> (imo - 1)) & 1) oc[no++] = imo;
      else vi[nvv++] = imo;
    }
    for (int k = 0; k < ni; ++k) L[k] = k;
    for (int k = 0; k < ne; ++k) L[ni + k] = ni + oc[k] - 1;
    int c = 0;
    for (int a = 0; a < ne; ++a)
      for (int b = a + 1; b < ne; ++b) { op0[c] = (short)a; op1[c] = (short)b; ++c; }
    nop_s = c;
    c = 0;
    for (int a = 0; a < nv; ++a)
      for (int b = a + 1; b < nv; ++b) { vp0[c] = (short)a; vp1[c] = (short)b; ++c; }
    nvp_s = c;
  }
  __syncthreads();
  const int64_t m = nmoc;
  auto Oi = [&](int a, int b, int c, int d) -> cplx { return O[a + m * (b + m * (c + m * (int64_t)d))]; };
  // diagonal (casmat.F90:57-104)
  const int nl = ni + ne;
  cplx part = mk(0.0, 0.0);
  for (int a0 = 0; a0 < nl; ++a0) {
    const int r = L[a0];
    for (int b0 = a0 + 1 + lane; b0 < nl; b0 += 64) {
      const int s = L[b0];
      const cplx x = Oi(r, r, s, s), y = Oi(r, s, s, r);
      part.x += 0.5 * x.x - 0.5 * y.x;
      part.y += 0.5 * x.y - 0.5 * y.y;
    }
  }
  part = wave_sum(part);
  if (lane == 0) {
    cplx d = mk(0.0, 0.0);
    for (int k = 0; k < ni; ++k) d = cadd(d, h[k + (int64_t)ldh * k]);
    for (int k = 0; k < ne; ++k) {
      const int g = ni + oc[k] - 1;
      d = cadd(d, h[g + (int64_t)ldh * g]);
    }
    mat[i + ndet * i] = cadd(d, cadd(part, cconj(part)));
  }
  // single excitations (casmat.F90:108-169)
  for (int idx = lane; idx < ne * nv; idx += 64) {
    const int r = oc[idx / nv], a = vi[idx % nv];
    int ph;
    const int64_t nw = one_e_exct(det, a, r, na, ph);
    const int64_t j = rank_cas[nw];
    if (j <= i) continue;
    const int ir = ni + r - 1, ia = ni + a - 1;
    cplx v = h[ir + (int64_t)ldh * ia];
    for (int k = 0; k < nl; ++k) {
      const int s = L[k];
      v = cadd(v, csub(Oi(ir, ia, s, s), Oi(ir, s, s, ia)));
    }
    if (ph & 1) v = mk(-v.x, -v.y);
    mat[i + ndet * j] = v;
    mat[j + ndet * i] = cconj(v);
  }
  // double excitations (casmat.F90:171-213)
  const int nop = nop_s, nvp = nvp_s;
  for (int64_t idx = lane; idx < (int64_t)nop * nvp; idx += 64) {
    const int po = (int)(idx / nvp), pv = (int)(idx % nvp);
    const int r = oc[op0[po]], s = oc[op1[po]], a = vi[vp0[pv]], b = vi[vp1[pv]];
    int ph1, ph2;
    const int64_t n1 = one_e_exct(det, a, r, na, ph1);
    const int64_t n2 = one_e_exct(n1, b, s, na, ph2);
    const int64_t j = rank_cas[n2];
    if (j <= i) continue;
    const int ir = ni + r - 1, is = ni + s - 1, ia = ni + a - 1, ib = ni + b - 1;
    cplx v = csub(Oi(ir, ia, is, ib), Oi(ir, ib, is, ia));
    if ((ph1 + ph2) & 1) v = mk(-v.x, -v.y);
    mat[i + ndet * j] = v;
    mat[j + ndet * i] = cconj(v);
  }
}

// fockcasci.F90:34-83 ; thread per element of the nmo x nmo matrix
__global__ void k_fock(int ni, int na, int ns, const cplx* __restrict__ h, int ldh, const cplx* __restrict__ O,
                       const cplx* __restrict__ F1, const cplx* __restrict__ F2, const cplx* __restrict__ D1,
                       cplx* fock) {
  const int nmoc = ni + na, nmo = nmoc + ns;
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= nmo || i > j) return;
  const int si = i < ni ? 0 : (i < nmoc ? 1 : 2), sj = j < ni ? 0 : (j < nmoc ? 1 : 2);
  if (si != sj) return;
  cplx f = h[i + (int64_t)ldh * j];
  if (si < 2) {
    const int64_t m = nmoc;
    auto Oi = [&](int a, int b, int c, int d) -> cplx { return O[a + m * (b + m * (c + m * (int64_t)d))]; };
    for (int k = 0; k < ni; ++k) f = cadd(f, csub(Oi(i, j, k, k), Oi(i, k, k, j)));
    for (int l = 0; l < na; ++l)
      for (int k = 0; k < na; ++k) {
        const cplx d = D1[k + na * l];
        cfma(f, d, csub(Oi(i, j, ni + k, ni + l), Oi(i, ni + l, ni + k, j)));
      }
  } else {
    const int64_t a = i - nmoc, b = j - nmoc, S = ns, m = nmoc;
    auto f1 = [&](int64_t x, int64_t y, int64_t k, int64_t l) { return F1[x + S * (y + S * (k + m * l))]; };
    auto f2 = [&](int64_t x, int64_t k, int64_t l, int64_t y) { return F2[x + S * (k + m * (l + m * y))]; };
    for (int k = 0; k < ni; ++k) f = cadd(f, csub(f1(a, b, k, k), f2(a, k, k, b)));
    for (int l = 0; l < na; ++l)
      for (int k = 0; k < na; ++k) {
        const cplx d = D1[k + na * l];
        cfma(f, d, csub(f1(a, b, ni + k, ni + l), f2(a, ni + l, ni + k, b)));
      }
  }
  fock[i + (int64_t)nmo * j] = f;
  if (i != j) fock[j + (int64_t)nmo * i] = cconj(f);
}

// fockivo.F90:46-78: the secondary-space Fock matrix of the IVO with the numh highest occupied
// spinors k = nmoc-numh .. nmoc-1 removed; thread per (a, b), a <= b, both triangles written
// (F_ba = conj(F_ab); the diagonal is conjugated as well, as the reference's hermitisation does)
__global__ void k_fockivo(int nmoc, int ns, int numh, int odd, const double* __restrict__ eps_sec,
                          const cplx* __restrict__ F1, const cplx* __restrict__ F2, cplx* __restrict__ fock) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (a >= ns || a > b) return;
  const int64_t S = ns, m = nmoc;
  cplx f = a == b ? mk(eps_sec[a], 0.0) : mk(0.0, 0.0);
  const double inv = 1.0 / (double)numh;
  for (int k = nmoc - numh; k < nmoc; ++k) {
    const cplx f1 = F1[a + S * (b + S * (k + m * k))];      // (ab|kk)
    const cplx f2 = F2[a + S * (k + m * (k + m * b))];      // (ak|kb)
    if (odd && k >= nmoc - 2) {
      f = csub(f, cscale(f1, 0.5 * inv));
      f = cadd(f, cscale(f2, 0.5 * inv));
    } else {
      f = cadd(f, cscale(f2, inv));
      f = csub(f, cscale(f1, inv));
    }
  }
  fock[b + S * a] = cconj(f);
  if (a != b) fock[a + S * b] = f;
}

__global__ void k_gather_block(int n, const int* ind, const cplx* fock, int ldf, cplx* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  out[i + (int64_t)n * j] = fock[ind[i] + (int64_t)ldf * ind[j]];
}

__global__ void k_scatter_block(int n, const int* ind, const cplx* V, const double* w, cplx* U, int ldu,
                                double* eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= n) return;
  U[ind[i] + (int64_t)ldu * ind[j]] = V[i + (int64_t)n * j];
  if (i == 0) eps[ind[j]] = w[j];
}

// Canonical gauge of Hermitian eigenvectors (the oracle's canonical_gauge).  The groups of
// (near-)degenerate eigenvalues: a group starts at i0 and takes every following j with
// |w_j - w_i0| <= 1e-8 max(1, |w_i0|); starts[g] is the first column of group g, starts[ng] = n.
__global__ void k_gauge_groups(int n, const double* w, int* starts) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int g = 0, i0 = 0;
  while (i0 < n) {
    starts[g++] = i0;
    int i1 = i0 + 1;
    while (i1 < n && fabs(w[i1] - w[i0]) <= 1.0e-8 * fmax(1.0, fabs(w[i0]))) ++i1;
    i0 = i1;
  }
  starts[g] = n;
  starts[n + 1] = g;
}

// One workgroup per group (blockIdx.x >= the group count: nothing to do); group g works in
// the columns starts[g] .. starts[g+1]-1 of V and of the scratch Ub (n x n).
__global__ __launch_bounds__(256) void k_gauge(int n, const int* starts, cplx* V, cplx* Ub) {
  __shared__ double dsh[4096];
  __shared__ double rmax[256];
  __shared__ int rmin[256];
  const int t = threadIdx.x;
  const int g = blockIdx.x;
  if (g >= starts[n + 1]) return;   // uniform over the workgroup
  {
    const int i0 = starts[g], i1 = starts[g + 1];
    const int m = i1 - i0;
    Ub += (int64_t)n * i0;
    for (int k = 0; k < m; ++k) {
      double lmax = -1.0;
      for (int r = t; r < n; r += 256) {
        double d = 0.0;
        for (int c = i0; c < i1; ++c) d += cabs2(V[r + (int64_t)n * c]);
        for (int l = 0; l < k; ++l) d -= cabs2(Ub[r + (int64_t)n * l]);
        dsh[r] = d;
        lmax = fmax(lmax, d);
      }
      rmax[t] = lmax;
      __syncthreads();
      for (int o = 128; o > 0; o >>= 1) {
        if (t < o) rmax[t] = fmax(rmax[t], rmax[t + o]);
        __syncthreads();
      }
      const double dmax = rmax[0];
      int lmin = 1 << 30;
      for (int r = t; r < n; r += 256)
        if (dsh[r] >= dmax - 1.0e-10 && r < lmin) lmin = r;
      rmin[t] = lmin;
      __syncthreads();
      for (int o = 128; o > 0; o >>= 1) {
        if (t < o) rmin[t] = min(rmin[t], rmin[t + o]);
        __syncthreads();
      }
      const int rk = rmin[0] < n ? rmin[0] : 0;   // NaN input: no row qualifies
      const double inv = 1.0 / sqrt(dsh[rk]);
      for (int r = t; r < n; r += 256) {
        cplx u = mk(0.0, 0.0);
        for (int c = i0; c < i1; ++c) cfma(u, V[r + (int64_t)n * c], cconj(V[rk + (int64_t)n * c]));
        for (int l = 0; l < k; ++l) {
          const cplx a = Ub[r + (int64_t)n * l], b = cconj(Ub[rk + (int64_t)n * l]);
          u = csub(u, cmul(a, b));
        }
        Ub[r + (int64_t)n * k] = cscale(u, inv);
      }
      __syncthreads();
    }
    for (int k = 0; k < m; ++k)
      for (int r = t; r < n; r += 256) V[r + (int64_t)n * (i0 + k)] = Ub[r + (int64_t)n * k];
  }
}

// ds(k, kt) for exactly NE electrons with every index static, so the NE x NE minor stays in
// registers: the same LU with partial pivoting, in the same order, as k_tracic_ds below.
template <int NE>
__global__ __launch_bounds__(128) void k_tracic_ds_reg(int na, int64_t ndet, const int64_t* dets,
                                                       const cplx* __restrict__ Ua, cplx* ds) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t kt = blockIdx.y;
  if (k >= ndet) return;
  int ok[NE], on[NE];
  {
    uint64_t a = (uint64_t)dets[k], b = (uint64_t)dets[kt];
#pragma unroll
    for (int r = 0; r < NE; ++r) {   // the r-th occupied orbital: lowest set bit, then clear it
      ok[r] = __ffsll((unsigned long long)a) - 1;
      on[r] = __ffsll((unsigned long long)b) - 1;
      a &= a - 1;
      b &= b - 1;
    }
  }
  cplx A[NE][NE];
#pragma unroll
  for (int r = 0; r < NE; ++r)
#pragma unroll
    for (int c = 0; c < NE; ++c) A[r][c] = Ua[ok[r] + (int64_t)na * on[c]];
  int swaps = 0;
#pragma unroll
  for (int c = 0; c < NE; ++c) {
    int piv = c;
    double best = fabs(A[c][c].x) + fabs(A[c][c].y);
#pragma unroll
    for (int r = c + 1; r < NE; ++r) {
      const double v = fabs(A[r][c].x) + fabs(A[r][c].y);
      if (v > best) { best = v; piv = r; }
    }
    if (piv != c) {
      ++swaps;
#pragma unroll
      for (int r = c + 1; r < NE; ++r)
        if (r == piv) {
#pragma unroll
          for (int q = 0; q < NE; ++q) { const cplx tmp = A[c][q]; A[c][q] = A[r][q]; A[r][q] = tmp; }
        }
    }
    const cplx d = A[c][c];
    const double den = cabs2(d);
    if (den == 0.0) continue;
    const cplx dinv = mk(d.x / den, -d.y / den);
#pragma unroll
    for (int r = c + 1; r < NE; ++r) {
      const cplx l = cmul(A[r][c], dinv);
#pragma unroll
      for (int q = c + 1; q < NE; ++q) A[r][q] = csub(A[r][q], cmul(l, A[c][q]));
    }
  }
  cplx det = mk(1.0, 0.0);
#pragma unroll
  for (int r = 0; r < NE; ++r) det = cmul(A[r][r], det);
  if (swaps & 1) det = mk(-det.x, -det.y);
  ds[k + ndet * kt] = det;
}

// any ne <= NE (the minor in private memory)
template <int NE>
__global__ void k_tracic_ds(int na, int ne, int64_t ndet, const int64_t* dets, const cplx* __restrict__ Ua,
                            cplx* ds) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t kt = blockIdx.y;
  if (k >= ndet) return;
  int ok[NE], on[NE];
  {
    const uint64_t a = (uint64_t)dets[k], b = (uint64_t)dets[kt];
    int x = 0, y = 0;
    for (int w = 0; w < na; ++w) {
      if ((a >> w) & 1ull) ok[x++] = w;
      if ((b >> w) & 1ull) on[y++] = w;
    }
  }
  cplx A[NE][NE];
  for (int r = 0; r < ne; ++r)
    for (int c = 0; c < ne; ++c) A[r][c] = Ua[ok[r] + (int64_t)na * on[c]];
  // LU with partial pivoting, pivot = first max of |re|+|im| (izamax)
  int swaps = 0;
  for (int c = 0; c < ne; ++c) {
    int piv = c;
    double best = fabs(A[c][c].x) + fabs(A[c][c].y);
    for (int r = c + 1; r < ne; ++r) {
      const double v = fabs(A[r][c].x) + fabs(A[r][c].y);
      if (v > best) { best = v; piv = r; }
    }
    if (piv != c) {
      ++swaps;
      for (int q = 0; q < ne; ++q) { const cplx tmp = A[c][q]; A[c][q] = A[piv][q]; A[piv][q] = tmp; }
    }
    const cplx d = A[c][c];
    const double den = cabs2(d);
    if (den == 0.0) continue;
    const cplx dinv = mk(d.x / den, -d.y / den);
    for (int r = c + 1; r < ne; ++r) {
      const cplx l = cmul(A[r][c], dinv);
      for (int q = c + 1; q < ne; ++q) A[r][q] = csub(A[r][q], cmul(l, A[c][q]));
    }
  }
  cplx det = mk(1.0, 0.0);
  for (int r = 0; r < ne; ++r) det = cmul(A[r][r], det);
  if (swaps & 1) det = mk(-det.x, -det.y);
  ds[k + ndet * kt] = det;
}

__global__ void k_calce0(int na, const double* eps_act, const cplx* D1, double* e0) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double e = 0.0;
    for (int w = 0; w < na; ++w) e += D1[w + na * w].x * eps_act[w];
    *e0 = e;
  }
}

struct Perm6 {
  int dims[6];
  int64_t src[6];
};
__global__ void k_permute6(Perm6 p, const cplx* __restrict__ in, cplx* __restrict__ out, int64_t total) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= total) return;
  int64_t r = o, s = 0;
  for (int a = 0; a < 6; ++a) {
    const int64_t q = r % p.dims[a];
    r /= p.dims[a];
    s += q * p.src[a];
  }
  out[o] = in[s];
}

struct Ext4 {
  int64_t ld[3];
  int off[4], n[4];
};
__global__ void k_extract4(Ext4 e, const cplx* __restrict__ in, cplx* __restrict__ out, int64_t total) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= total) return;
  int64_t r = o;
  const int i = (int)(r % e.n[0]); r /= e.n[0];
  const int j = (int)(r % e.n[1]); r /= e.n[1];
  const int k = (int)(r % e.n[2]); r /= e.n[2];
  const int l = (int)r;
  out[o] = in[(e.off[0] + i) + e.ld[0] * ((e.off[1] + j) + e.ld[1] * ((e.off[2] + k) + e.ld[2] * (int64_t)(e.off[3] + l)))];
}

__global__ void k_fill_zero(cplx* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = mk(0.0, 0.0);
}

__global__ void k_flag_clear(int* flag) { *flag = 0; }

// flag = 1 when any of the n doubles is not finite (each workgroup that finds one stores 1)
__global__ __launch_bounds__(256) void k_nonfinite(const double* __restrict__ p, int64_t n, int* __restrict__ flag) {
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    bad = bad || !isfinite(p[i]);
  if (__syncthreads_or(bad) && threadIdx.x == 0) *flag = 1;
}

__global__ void k_scale_cols_tail(int64_t n, int64_t dimm, const cplx* vs, const double* ws, int inverse,
                                  cplx* uc) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (r >= n) return;
  const int64_t src = n - dimm + c;
  uc[r + n * c] = cscale(vs[r + n * src], inverse ? 1.0 / ws[src] : pow(ws[src], -0.5));
}

// C(n, k), exact in 64-bit for n <= 62
__device__ inline int64_t binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  int64_t r = 1;
  for (int i = 0; i < k; ++i) r = r * (n - i) / (i + 1);
  return r;
}

// strings[r] = the r-th k-bit string below 2^na in ascending order (colex unranking:
// rank = sum_j C(p_j, j) over the set bit positions p_1 < ... < p_k)
__global__ void k_unrank(int na, int k, int64_t m, int64_t* strings) {
  const int64_t r0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r0 >= m) return;
  int64_t r = r0, d = 0;
  int p = na - 1;
  for (int j = k; j >= 1; --j) {
    while (binom(p, j) > r) --p;
    d |= 1ll << p;
    r -= binom(p, j);
    --p;
  }
  strings[r0] = d;
}

// Phi[r + m*(t*na*na + u*na + v)] = <r| a+_t E_uv |0> (create) or <r| a_t E_uv |0>, gathered
// from Y1[:, u + na*v] = E_uv|0>.  The ladder phase counts the occupied spinors above t,
// the convention of one_e_exct, so that a+_t a_x = E_tx holds with the tables of apply_E.
__global__ void k_phi(int na, int create, int64_t m, const int64_t* __restrict__ strings, int64_t nf,
                      const int* __restrict__ rank_full, const cplx* __restrict__ Y1, cplx* __restrict__ Phi) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const int t = blockIdx.y, uv = blockIdx.z;
  const int u = uv / na, v = uv % na;
  const uint64_t d = (uint64_t)strings[r], bit = 1ull << t;
  cplx val = mk(0.0, 0.0);
  if (((d & bit) != 0) == (create != 0)) {
    const int k = rank_full[d ^ bit];
    val = Y1[(int64_t)k + nf * (u + na * v)];
    if (__popcll(d >> (t + 1)) & 1) val = mk(-val.x, -val.y);
  }
  Phi[r + m * ((int64_t)t * na * na + uv)] = val;
}

// The antisymmetric reduction of the A factor, stored adjoint: row k of Psi is the tuple
// tup[k] = t*na*na + u*na + v (t < u) of Phi, Psi[k + nred*r] = conj(<r| a+_t E_uv |0>) over
// the (ne+1)-electron strings r, and Psif[k + nred*r] = f[r] Psi[k + nred*r] with f[r] =
// sum_w eps_w n_w(r).  Threads run along k, so the stores of a row r are contiguous; r
// strides over the grid's y dimension.
__global__ void k_phi_adj(int na, int64_t m, const int64_t* __restrict__ strings, int64_t nf,
                          const int* __restrict__ rank_full, const cplx* __restrict__ Y1,
                          const double* __restrict__ eps_act, int64_t nred, const int* __restrict__ tup,
                          cplx* __restrict__ Psi, cplx* __restrict__ Psif) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nred) return;
  const int idx = tup[k];
  const int t = idx / (na * na), u = (idx / na) % na, v = idx % na;
  const uint64_t bit = 1ull << t;
  for (int64_t r = blockIdx.y; r < m; r += gridDim.y) {
    const uint64_t d = (uint64_t)strings[r];
    double f = 0.0;
    for (int w = 0; w < na; ++w)
      if ((d >> w) & 1ull) f += eps_act[w];
    cplx val = mk(0.0, 0.0);
    if (d & bit) {
      val = Y1[(int64_t)rank_full[d ^ bit] + nf * (u + na * v)];
      if (__popcll(d >> (t + 1)) & 1) val = mk(-val.x, -val.y);
    }
    val = cconj(val);
    Psi[k + nred * r] = val;
    Psif[k + nred * r] = mk(f * val.x, f * val.y);
  }
}

// pair index p = w (w - 1) / 2 + s of s < w, and its inverse
__device__ inline int thin_pair(int a, int b) {
  const int s = a < b ? a : b, w = a < b ? b : a;
  return w * (w - 1) / 2 + s;
}
__device__ inline void thin_unpair(int p, int& s, int& w) {
  w = (int)((1.0 + sqrt(1.0 + 8.0 * p)) * 0.5);
  while (w * (w - 1) / 2 > p) --w;
  while ((w + 1) * w / 2 <= p) ++w;
  s = p - w * (w - 1) / 2;
}

// A_v[r] = <r| a_v |0> from the C factor: sum_t a_t a+_t a_v |0> = (na - ne + 1) a_v |0> on the
// (ne-1)-electron strings, Phi[r + m (t na na + u na + v)] = <r| a_t a+_u a_v |0>.
__global__ void k_thin_avec(int na, int ne, int64_t m, const cplx* __restrict__ Phi, cplx* __restrict__ Av) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const int v = blockIdx.y;
  const int64_t na2 = (int64_t)na * na;
  cplx acc = mk(0.0, 0.0);
  for (int t = 0; t < na; ++t) acc = cadd(acc, Phi[r + m * (t * na2 + (int64_t)t * na + v)]);
  const double inv = 1.0 / (double)(na - ne + 1);
  Av[r + m * v] = mk(acc.x * inv, acc.y * inv);
}

// Thin factor of the C metric G = Phi Phi^H, stored adjoint: with Psi_{u,(s,w)} = a+_u a_s a_w |0>
// (s < w), Phi = Psi M for a fixed real M, and M M^T = 2 I + kappa sum_v sigma_v sigma_v^T,
// kappa = (na - 2 (ne-1)) / (ne-1)^2, over disjoint blocks (the columns (u, (s, w)) with u in
// {s, w}, v the other index, sigma = +1 for u < v, -1 for u > v).  Its square root gives
// Psi' = Psi (M M^T)^(1/2) with Psi' Psi'^H = G:
//   Psi'_{u,(s,w)} = sqrt2 (Phi(w,u,s) - [w = u] A_s) + [u = s] beta (ne-1) A_w - [u = w] beta (ne-1) A_s,
//   beta = (sqrt(2 + kappa (na-1)) - sqrt2) / (na-1),
// since Phi(w,u,s) = [w = u] a_s|0> + Psi_{u,(s,w)} and sum_x sigma_x Psi_{x,(x,v)} = (ne-1) a_v|0>.
// Column j = u * npair + p, pair p = w (w - 1) / 2 + s (s < w); PsiH[j + nred r] = conj(Psi'[r, j]).
__global__ void k_thin_c(int na, int ne, int64_t m, const cplx* __restrict__ Phi, const cplx* __restrict__ Av,
                         int64_t nred, cplx* __restrict__ PsiH) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nred) return;
  const int npair = na * (na - 1) / 2;
  const int u = (int)(j / npair);
  int sidx, w;
  thin_unpair((int)(j % npair), sidx, w);
  const int64_t na2 = (int64_t)na * na;
  const double r2 = sqrt(2.0);
  const double kappa = (double)(na - 2 * (ne - 1)) / ((double)(ne - 1) * (ne - 1));
  const double beta = (sqrt(2.0 + kappa * (na - 1)) - r2) / (double)(na - 1) * (double)(ne - 1);
  const int64_t col = (int64_t)w * na2 + (int64_t)u * na + sidx;
  for (int64_t r = blockIdx.y; r < m; r += gridDim.y) {
    cplx x = Phi[r + m * col];
    if (w == u) x = csub(x, Av[r + m * sidx]);
    x = mk(r2 * x.x, r2 * x.y);
    if (u == sidx) cfma(x, mk(beta, 0.0), Av[r + m * w]);
    if (u == w) cfma(x, mk(-beta, 0.0), Av[r + m * sidx]);
    PsiH[j + nred * r] = cconj(x);
  }
}

// MV = M V for the thin C factor (Phi = Psi M), V (na^3 x npert, row t na^2 + u na + v):
//   (M V)_{u,(s,w)} = V(w,u,s) - V(s,u,w) + [u in {s,w}] sigma/(ne-1) sum_t V(t,t,v),
// v the other pair index, sigma = +1 for u = s (u < v), -1 for u = w.
__global__ void k_thin_mv(int na, int ne, int64_t npert, const cplx* __restrict__ V, int64_t nred,
                          cplx* __restrict__ MV) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nred) return;
  const int npair = na * (na - 1) / 2;
  const int u = (int)(j / npair);
  int sidx, w;
  thin_unpair((int)(j % npair), sidx, w);
  const int64_t na2 = (int64_t)na * na, na3 = na2 * na;
  const int sg = u == sidx ? 1 : (u == w ? -1 : 0);
  const int v = u == sidx ? w : sidx;
  const double inv = 1.0 / (double)(ne - 1);
  for (int64_t p = blockIdx.y; p < npert; p += gridDim.y) {
    const cplx* Vp = V + na3 * p;
    cplx x = csub(Vp[(int64_t)w * na2 + (int64_t)u * na + sidx], Vp[(int64_t)sidx * na2 + (int64_t)u * na + w]);
    if (sg) {
      cplx tr = mk(0.0, 0.0);
      for (int t = 0; t < na; ++t) tr = cadd(tr, Vp[(int64_t)t * na2 + (int64_t)t * na + v]);
      cfma(x, mk(sg * inv, 0.0), tr);
    }
    MV[j + nred * p] = x;
  }
}

// W = (M M^T)^(-1/2) M V: 2^(-1/2) on the columns outside the blocks; on block v
// (2 I + kappa sigma sigma^T)^(-1/2) = 2^(-1/2) I + gamma sigma sigma^T,
// gamma = ((2 + kappa (na-1))^(-1/2) - 2^(-1/2)) / (na-1).
__global__ void k_thin_w(int na, int ne, int64_t npert, const cplx* __restrict__ MV, int64_t nred,
                         cplx* __restrict__ W) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nred) return;
  const int npair = na * (na - 1) / 2;
  const int u = (int)(j / npair);
  int sidx, w;
  thin_unpair((int)(j % npair), sidx, w);
  const int sg = u == sidx ? 1 : (u == w ? -1 : 0);
  const int v = u == sidx ? w : sidx;
  const double kappa = (double)(na - 2 * (ne - 1)) / ((double)(ne - 1) * (ne - 1));
  const double r2 = 1.0 / sqrt(2.0);
  const double gam = (1.0 / sqrt(2.0 + kappa * (na - 1)) - r2) / (double)(na - 1);
  for (int64_t p = blockIdx.y; p < npert; p += gridDim.y) {
    const cplx* M = MV + nred * p;
    cplx x = mk(r2 * M[j].x, r2 * M[j].y);
    if (sg) {
      cplx b = mk(0.0, 0.0);
      for (int y = 0; y < na; ++y) {
        if (y == v) continue;
        const cplx e = M[(int64_t)y * npair + thin_pair(y, v)];
        b = y < v ? cadd(b, e) : csub(b, e);
      }
      cfma(x, mk(gam * sg, 0.0), b);
    }
    W[j + nred * p] = x;
  }
}

// f[r] = sum_w eps_w n_w(strings[r])
__global__ void k_fdet(int na, int64_t m, const int64_t* strings, const double* eps_act, double* f) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const uint64_t d = (uint64_t)strings[r];
  double s = 0.0;
  for (int w = 0; w < na; ++w)
    if ((d >> w) & 1ull) s += eps_act[w];
  f[r] = s;
}

// ---- block Davidson helpers
__global__ void k_diag_real(int64_t n, const cplx* H, double* d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = H[i + n * i].x;
}

__global__ void k_unit_cols(int64_t n, int ncol, const int* rows, cplx* V) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i < n) V[i + n * j] = mk(i == rows[j] ? 1.0 : 0.0, 0.0);
}

// R[:, j] = AX[:, j] - theta_j X[:, j]
__global__ void k_ritz_residual(int64_t n, const cplx* AX, const cplx* X, const double* theta, cplx* R) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i < n) R[i + n * j] = csub(AX[i + n * j], cscale(X[i + n * j], theta[j]));
}

// T[:, c] = R[:, sel_c] / (theta_{sel_c} - d), denominators kept at least 1e-8 in magnitude
__global__ void k_precond(int64_t n, const int* sel, const cplx* R, const double* theta, const double* d, cplx* T) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y, j = sel[c];
  if (i >= n) return;
  double den = theta[j] - d[i];
  if (fabs(den) < 1e-8) den = den < 0 ? -1e-8 : 1e-8;
  T[i + n * c] = cscale(R[i + n * j], 1.0 / den);
}

// Ritz vectors, residuals and preconditioned corrections of one Davidson iteration, one thread
// per (row i, column j): x = sum_k V[i, k] Y[k, j] and ax = sum_k W[i, k] Y[k, j] in the order
// k = 0..mm-1, r = ax - theta_j x, t = r / (theta_j - d_i) with the denominator kept at least
// 1e-8 in magnitude (k_precond's rule); Y's column is uniform over the workgroup
__global__ __launch_bounds__(256) void k_ritz_fused(int64_t n, int mm, const cplx* __restrict__ V,
                                                    const cplx* __restrict__ W, const cplx* __restrict__ Y,
                                                    const double* __restrict__ theta, const double* __restrict__ d,
                                                    cplx* __restrict__ X, cplx* __restrict__ AX,
                                                    cplx* __restrict__ R, cplx* __restrict__ T) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  const cplx* y = Y + (int64_t)mm * j;
  cplx x = mk(0.0, 0.0), ax = mk(0.0, 0.0);
  for (int k = 0; k < mm; ++k) {
    const cplx yk = y[k];
    cfma(x, V[i + n * k], yk);
    cfma(ax, W[i + n * k], yk);
  }
  const double th = theta[j];
  const cplx r = csub(ax, cscale(x, th));
  double den = th - d[i];
  if (fabs(den) < 1e-8) den = den < 0 ? -1e-8 : 1e-8;
  X[i + n * j] = x;
  AX[i + n * j] = ax;
  R[i + n * j] = r;
  T[i + n * j] = cscale(r, 1.0 / den);
}

// one workgroup per column: T[:, c] /= |T[:, c]| (deterministic tree reduction)
__global__ __launch_bounds__(256) void k_col_normalize(int64_t n, cplx* T) {
  __shared__ double sh[256];
  const int64_t c = blockIdx.x;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += cabs2(T[i + n * c]);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  const double nrm = sqrt(sh[0]);
  const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) T[i + n * c] = cscale(T[i + n * c], inv);
}

__global__ void k_scale_rows(int64_t m, int64_t ncol, const double* f, const cplx* in, cplx* out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (r >= m) return;
  out[r + m * c] = cscale(in[r + m * c], f[r]);
}

}  // namespace

void enumerate_dets(hipStream_t s, const DetParams& p, int* ff, int* fc, int* rf, int* rc, int64_t* full_dets,
                    int64_t* cas_dets, int* cas2full, int64_t* n_full, int64_t* n_cas, int* counts_dev) {
  DC2_CHECK(p.nact >= 1 && p.nact <= 24, 1, "nact must be in 1..24");
  DC2_CHECK(p.nelec >= 1 && p.nelec <= p.nact, 1, "nelec must be in 1..nact");
  const int64_t N = 1ll << p.nact;
  hipLaunchKernelGGL(k_det_flags, dim3(cdiv(N, 256)), dim3(256), 0, s, p, N, ff, fc);
  hipLaunchKernelGGL(k_scan2, dim3(1), dim3(1024), 0, s, N, ff, fc, rf, rc, counts_dev);
  int counts[2];
  DC2_HIP(hipMemcpyAsync(counts, counts_dev, sizeof(counts), hipMemcpyDeviceToHost, s));
  DC2_HIP(hipStreamSynchronize(s));
  *n_full = counts[0];
  *n_cas = counts[1];
  if (full_dets)
    hipLaunchKernelGGL(k_det_scatter, dim3(cdiv(N, 256)), dim3(256), 0, s, N, rf, rc, full_dets, cas_dets, cas2full);
  DC2_HIP(hipGetLastError());
}

void build_exc_table(hipStream_t s, int na, int64_t nf, const int64_t* full_dets, const int* rank_full, int* tab) {
  hipLaunchKernelGGL(k_exc_table, dim3(cdiv(nf, 256), na * na), dim3(256), 0, s, na, nf, full_dets, rank_full, tab);
  DC2_HIP(hipGetLastError());
}

void apply_E(hipStream_t s, int na, int64_t nf, int64_t nx, const int* tab, const cplx* X, cplx* Y) {
  DC2_CHECK(nx < 65536, 1, "apply_E: too many right-hand sides");
  hipLaunchKernelGGL(k_apply_E, dim3(cdiv(nf, 256), na * na, (unsigned)nx), dim3(256), 0, s, na, nf, tab, X, Y);
  DC2_HIP(hipGetLastError());
}

void embed_ci(hipStream_t s, int na, int64_t nf, int64_t ncas, const int* cas2full, const cplx* c_cas,
              const int64_t* full_dets, const double* eps_act, cplx* c_full, cplx* cF) {
  hipLaunchKernelGGL(k_fill_zero, dim3(cdiv(nf, 256)), dim3(256), 0, s, c_full, nf);
  hipLaunchKernelGGL(k_embed, dim3(cdiv(ncas, 256)), dim3(256), 0, s, na, ncas, cas2full, c_cas, c_full);
  hipLaunchKernelGGL(k_fvec, dim3(cdiv(nf, 256)), dim3(256), 0, s, na, nf, full_dets, eps_act, c_full, cF);
  DC2_HIP(hipGetLastError());
}

void casmat(hipStream_t s, int ni, int na, int ne, int64_t ndet, const int64_t* cas_dets, const int* rank_cas,
            const cplx* O, int nmoc, const cplx* h, int ldh, cplx* mat) {
  DC2_CHECK(na <= 32 && ni + ne <= 96, 1, "casmat: nact <= 32 and ninact + nelec <= 96 are supported");
  hipLaunchKernelGGL(k_fill_zero, dim3(cdiv(ndet * ndet, 256)), dim3(256), 0, s, mat, ndet * ndet);
  hipLaunchKernelGGL(k_casmat, dim3((unsigned)ndet), dim3(64), 0, s, ni, na, ne, ndet, cas_dets, rank_cas, O, nmoc,
                     h, ldh, mat);
  DC2_HIP(hipGetLastError());
}

void fock_build(hipStream_t s, int ni, int na, int ns, const cplx* h, int ldh, const cplx* O, const cplx* F1,
                const cplx* F2, const cplx* D1, cplx* fock) {
  const int nmo = ni + na + ns;
  hipLaunchKernelGGL(k_fill_zero, dim3(cdiv((int64_t)nmo * nmo, 256)), dim3(256), 0, s, fock, (int64_t)nmo * nmo);
  hipLaunchKernelGGL(k_fock, dim3(cdiv(nmo, 128), nmo), dim3(128), 0, s, ni, na, ns, h, ldh, O, F1, F2, D1, fock);
  DC2_HIP(hipGetLastError());
}

void fockivo_build(hipStream_t s, int nmoc, int ns, int numh, bool odd_nelec, const double* eps_sec, const cplx* F1,
                   const cplx* F2, cplx* fock) {
  if (ns <= 0) return;
  hipLaunchKernelGGL(k_fockivo, dim3(cdiv(ns, 128), ns), dim3(128), 0, s, nmoc, ns, numh, odd_nelec ? 1 : 0, eps_sec,
                     F1, F2, fock);
  DC2_HIP(hipGetLastError());
}

void gather_block(hipStream_t s, int n, const int* ind, const cplx* fock, int ldf, cplx* out) {
  hipLaunchKernelGGL(k_gather_block, dim3(cdiv(n, 128), n), dim3(128), 0, s, n, ind, fock, ldf, out);
  DC2_HIP(hipGetLastError());
}

void scatter_block(hipStream_t s, int n, const int* ind, const cplx* V, const double* w, cplx* U, int ldu,
                   double* eps) {
  hipLaunchKernelGGL(k_scatter_block, dim3(cdiv(n, 128), n), dim3(128), 0, s, n, ind, V, w, U, ldu, eps);
  DC2_HIP(hipGetLastError());
}

void canonical_gauge(hipStream_t s, int n, const double* w, cplx* V, cplx* scratch, int* starts) {
  DC2_CHECK(n <= 4096, 1, "canonical_gauge: block larger than 4096");
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gauge_groups, dim3(1), dim3(64), 0, s, n, w, starts);
  hipLaunchKernelGGL(k_gauge, dim3(n), dim3(256), 0, s, n, starts, V, scratch);
  DC2_HIP(hipGetLastError());
}

void tracic_ds(hipStream_t s, int na, int ne, int64_t ndet, const int64_t* cas_dets, const cplx* Uact, cplx* ds) {
  DC2_CHECK(ndet < 65536, 1, "tracic: more than 65535 determinants");
  dim3 g(cdiv(ndet, 128), (unsigned)ndet);
  switch (ne) {   // exact sizes with the minor in registers
    case 1: hipLaunchKernelGGL(k_tracic_ds_reg<1>, g, dim3(128), 0, s, na, ndet, cas_dets, Uact, ds); break;
    case 2: hipLaunchKernelGGL(k_tracic_ds_reg<2>, g, dim3(128), 0, s, na, ndet, cas_dets, Uact, ds); break;
    case 3: hipLaunchKernelGGL(k_tracic_ds_reg<3>, g, dim3(128), 0, s, na, ndet, cas_dets, Uact, ds); break;
    case 4: hipLaunchKernelGGL(k_tracic_ds_reg<4>, g, dim3(128), 0, s, na, ndet, cas_dets, Uact, ds); break;
    case 5: hipLaunchKernelGGL(k_tracic_ds_reg<5>, g, dim3(128), 0, s, na, ndet, cas_dets, Uact, ds); break;
    case 6: hipLaunchKernelGGL(k_tracic_ds_reg<6>, g, dim3(128), 0, s, na, ndet, cas_dets, Uact, ds); break;
    case 7: hipLaunchKernelGGL(k_tracic_ds_reg<7>, g, dim3(128), 0, s, na, ndet, cas_dets, Uact, ds); break;
    case 8: hipLaunchKernelGGL(k_tracic_ds_reg<8>, g, dim3(128), 0, s, na, ndet, cas_dets, Uact, ds); break;
    default: break;
  }
  if (ne >= 1 && ne <= 8) {
    DC2_HIP(hipGetLastError());
    return;
  }
  if (ne <= 4) hipLaunchKernelGGL(k_tracic_ds<4>, g, dim3(128), 0, s, na, ne, ndet, cas_dets, Uact, ds);
  else if (ne <= 8) hipLaunchKernelGGL(k_tracic_ds<8>, g, dim3(128), 0, s, na, ne, ndet, cas_dets, Uact, ds);
  else if (ne <= 16) hipLaunchKernelGGL(k_tracic_ds<16>, g, dim3(128), 0, s, na, ne, ndet, cas_dets, Uact, ds);
  else throw Error(1, "tracic: nelec > 16 not supported");
  DC2_HIP(hipGetLastError());
}

void calce0(hipStream_t s, int na, const double* eps_act, const cplx* D1, double* e0) {
  hipLaunchKernelGGL(k_calce0, dim3(1), dim3(64), 0, s, na, eps_act, D1, e0);
  DC2_HIP(hipGetLastError());
}

void permute6(hipStream_t s, const int* dims_out, const int* src_stride, const cplx* in, cplx* out, int64_t total) {
  Perm6 p;
  for (int a = 0; a < 6; ++a) { p.dims[a] = dims_out[a]; p.src[a] = src_stride[a]; }
  hipLaunchKernelGGL(k_permute6, dim3(cdiv(total, 256)), dim3(256), 0, s, p, in, out, total);
  DC2_HIP(hipGetLastError());
}

void extract4(hipStream_t s, const cplx* in, const int64_t* ld_in, const int* off, const int* n, cplx* out) {
  Ext4 e;
  for (int a = 0; a < 3; ++a) e.ld[a] = ld_in[a];
  for (int a = 0; a < 4; ++a) { e.off[a] = off[a]; e.n[a] = n[a]; }
  const int64_t total = (int64_t)n[0] * n[1] * n[2] * n[3];
  if (total == 0) return;
  hipLaunchKernelGGL(k_extract4, dim3(cdiv(total, 256)), dim3(256), 0, s, e, in, out, total);
  DC2_HIP(hipGetLastError());
}

void fill_zero(hipStream_t s, cplx* p, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_zero, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n);
  DC2_HIP(hipGetLastError());
}

bool all_finite_dev(hipStream_t s, const double* p, int64_t n, int* flag_dev, int* flag_host) {
  hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(1), 0, s, flag_dev);
  if (n > 0)
    hipLaunchKernelGGL(k_nonfinite, dim3((unsigned)std::min<int64_t>(cdiv(n, 256 * 8), 1024)), dim3(256), 0, s, p, n,
                       flag_dev);
  DC2_HIP(hipGetLastError());
  DC2_HIP(hipMemcpyAsync(flag_host, flag_dev, sizeof(int), hipMemcpyDeviceToHost, s));
  DC2_HIP(hipStreamSynchronize(s));
  return *flag_host == 0;
}

void scale_cols_tail(hipStream_t s, int64_t n, int64_t dimm, const cplx* vs, const double* ws, cplx* uc,
                     bool inverse) {
  if (dimm <= 0 || n <= 0) return;
  DC2_CHECK(dimm < 65536, 1, "scale_cols_tail: too many columns for one grid dimension");
  hipLaunchKernelGGL(k_scale_cols_tail, dim3(cdiv(n, 256), (unsigned)dimm), dim3(256), 0, s, n, dimm, vs, ws,
                     inverse ? 1 : 0, uc);
  DC2_HIP(hipGetLastError());
}

int64_t binom_host(int n, int k) {
  if (k < 0 || k > n) return 0;
  int64_t r = 1;
  for (int i = 0; i < k; ++i) r = r * (n - i) / (i + 1);
  return r;
}

void build_phi(hipStream_t s, int na, int ne, bool create, int64_t nf, const int* rank_full, const cplx* Y1,
               const double* eps_act, int64_t m, int64_t* strings, cplx* Phi, double* fdiag) {
  DC2_CHECK(m == binom_host(na, create ? ne + 1 : ne - 1), 1, "build_phi: wrong string count");
  if (m == 0) return;
  DC2_CHECK(na >= 1 && na <= 24 && (int64_t)na * na < 65536, 1, "build_phi: nact out of range");
  hipLaunchKernelGGL(k_unrank, dim3(cdiv(m, 256)), dim3(256), 0, s, na, create ? ne + 1 : ne - 1, m, strings);
  hipLaunchKernelGGL(k_phi, dim3(cdiv(m, 256), na, na * na), dim3(256), 0, s, na, create ? 1 : 0, m, strings, nf,
                     rank_full, Y1, Phi);
  hipLaunchKernelGGL(k_fdet, dim3(cdiv(m, 256)), dim3(256), 0, s, na, m, strings, eps_act, fdiag);
  DC2_HIP(hipGetLastError());
}

void build_thin_c(hipStream_t s, int na, int ne, int64_t m, const cplx* Phi, cplx* Av, cplx* PsiH) {
  if (m == 0) return;
  DC2_CHECK(na >= 2 && ne >= 2 && ne <= na, 1, "build_thin_c: needs 2 <= nelec <= nact");
  const int64_t nred = (int64_t)na * na * (na - 1) / 2;
  hipLaunchKernelGGL(k_thin_avec, dim3(cdiv(m, 256), na), dim3(256), 0, s, na, ne, m, Phi, Av);
  hipLaunchKernelGGL(k_thin_c, dim3(cdiv(nred, 256), (unsigned)std::min<int64_t>(m, 65535)), dim3(256), 0, s, na, ne,
                     m, Phi, Av, nred, PsiH);
  DC2_HIP(hipGetLastError());
}

void thin_c_w(hipStream_t s, int na, int ne, int64_t npert, const cplx* V, cplx* MV, cplx* W) {
  const int64_t nred = (int64_t)na * na * (na - 1) / 2;
  if (npert <= 0 || nred == 0) return;
  const dim3 grid((unsigned)cdiv(nred, 256), (unsigned)std::min<int64_t>(npert, 65535));
  hipLaunchKernelGGL(k_thin_mv, grid, dim3(256), 0, s, na, ne, npert, V, nred, MV);
  hipLaunchKernelGGL(k_thin_w, grid, dim3(256), 0, s, na, ne, npert, MV, nred, W);
  DC2_HIP(hipGetLastError());
}

void build_phi_adj(hipStream_t s, int na, int ne, int64_t nf, const int* rank_full, const cplx* Y1,
                   const double* eps_act, int64_t m, int64_t* strings, int64_t nred, const int* tup, cplx* Psi,
                   cplx* Psif) {
  DC2_CHECK(m == binom_host(na, ne + 1), 1, "build_phi_adj: wrong string count");
  if (m == 0 || nred == 0) return;
  DC2_CHECK(na >= 2 && na <= 24, 1, "build_phi_adj: nact out of range");
  DC2_CHECK(nred == (int64_t)na * na * (na - 1) / 2, 1, "build_phi_adj: wrong tuple count");
  hipLaunchKernelGGL(k_unrank, dim3(cdiv(m, 256)), dim3(256), 0, s, na, ne + 1, m, strings);
  hipLaunchKernelGGL(k_phi_adj, dim3(cdiv(nred, 256), (unsigned)std::min<int64_t>(m, 65535)), dim3(256), 0, s, na, m,
                     strings, nf, rank_full, Y1, eps_act, nred, tup, Psi, Psif);
  DC2_HIP(hipGetLastError());
}

void davidson_diag(hipStream_t s, int64_t n, const cplx* H, double* d) {
  hipLaunchKernelGGL(k_diag_real, dim3(cdiv(n, 256)), dim3(256), 0, s, n, H, d);
  DC2_HIP(hipGetLastError());
}
void davidson_unit(hipStream_t s, int64_t n, int ncol, const int* rows, cplx* V) {
  hipLaunchKernelGGL(k_unit_cols, dim3(cdiv(n, 256), ncol), dim3(256), 0, s, n, ncol, rows, V);
  DC2_HIP(hipGetLastError());
}
void davidson_residual(hipStream_t s, int64_t n, int ncol, const cplx* AX, const cplx* X, const double* theta,
                       cplx* R) {
  hipLaunchKernelGGL(k_ritz_residual, dim3(cdiv(n, 256), ncol), dim3(256), 0, s, n, AX, X, theta, R);
  DC2_HIP(hipGetLastError());
}
void davidson_precond(hipStream_t s, int64_t n, int ncol, const int* sel, const cplx* R, const double* theta,
                      const double* d, cplx* T) {
  hipLaunchKernelGGL(k_precond, dim3(cdiv(n, 256), ncol), dim3(256), 0, s, n, sel, R, theta, d, T);
  DC2_HIP(hipGetLastError());
}

void davidson_ritz(hipStream_t s, int64_t n, int ncol, int mm, const cplx* V, const cplx* W, const cplx* Sy,
                   const double* theta, const double* d, cplx* X, cplx* AX, cplx* R, cplx* T) {
  DC2_CHECK(mm >= 1 && mm <= 64 && ncol >= 1 && ncol < 65536, 1, "davidson_ritz: subspace out of range");
  hipLaunchKernelGGL(k_ritz_fused, dim3(cdiv(n, 256), ncol), dim3(256), 0, s, n, mm, V, W, Sy, theta, d, X, AX, R,
                     T);
  DC2_HIP(hipGetLastError());
}

// Orthonormal combinations of a Davidson correction block of four columns from its Gram
// matrix G = T^H T (4 x 4): a Cholesky factorisation G = R^H R that skips the columns whose
// remaining squared norm is at most 1e-12 (their directions already lie in the block), then
// Sc = R^-1 on the kept columns, written first, zero columns after them, and the kept count in
// nk[0].  One thread, every index static.  This is the block orthonormalisation the host did
// between two round trips, so that an iteration needs one host round trip instead of two.
__global__ __launch_bounds__(64) void k_block_chol4(const cplx* __restrict__ G, cplx* __restrict__ Sc,
                                                    int* __restrict__ nk) {
  constexpr int B = 4;
  if (threadIdx.x != 0) return;
  cplx A[B][B], R[B][B];
  bool keep[B];
#pragma unroll
  for (int j = 0; j < B; ++j)
#pragma unroll
    for (int i = 0; i < B; ++i) {
      A[i][j] = i <= j ? G[i + B * j] : cconj(G[j + B * i]);
      R[i][j] = mk(0.0, 0.0);
    }
  // R[j][k] for kept rows j < k: (A[j][k] - sum_{i<j kept} conj(R[i][j]) R[i][k]) / R[j][j]
#pragma unroll
  for (int k = 0; k < B; ++k) {
    double d = A[k][k].x;
#pragma unroll
    for (int j = 0; j < k; ++j) {
      if (!keep[j]) continue;
      cplx x = A[j][k];
#pragma unroll
      for (int i = 0; i < j; ++i)
        if (keep[i]) x = csub(x, cmul(cconj(R[i][j]), R[i][k]));
      x = mk(x.x / R[j][j].x, x.y / R[j][j].x);
      R[j][k] = x;
      d -= cabs2(x);
    }
    keep[k] = d > 1e-12;
    R[k][k] = mk(keep[k] ? sqrt(d) : 1.0, 0.0);
  }
  // X = R^-1 (upper triangular on the kept indices): X[k][k] = 1/R[k][k],
  // X[i][k] = -(sum_{i<=j<k, kept} X[i][j] R[j][k]) / R[k][k]
  cplx X[B][B];
#pragma unroll
  for (int k = 0; k < B; ++k)
#pragma unroll
    for (int i = 0; i < B; ++i) X[i][k] = mk(0.0, 0.0);
#pragma unroll
  for (int k = 0; k < B; ++k) {
    if (!keep[k]) continue;
    const double inv = 1.0 / R[k][k].x;
    X[k][k] = mk(inv, 0.0);
#pragma unroll
    for (int i = 0; i < k; ++i) {
      if (!keep[i]) continue;
      cplx sum = mk(0.0, 0.0);
#pragma unroll
      for (int j = i; j < k; ++j)
        if (keep[j]) cfma(sum, X[i][j], R[j][k]);
      X[i][k] = mk(-sum.x * inv, -sum.y * inv);
    }
  }
  int kept = 0;
#pragma unroll
  for (int k = 0; k < B; ++k) {
    if (keep[k]) {
#pragma unroll
      for (int i = 0; i < B; ++i) Sc[i + B * kept] = X[i][k];
      ++kept;
    }
  }
  for (int k = kept; k < B; ++k)
#pragma unroll
    for (int i = 0; i < B; ++i) Sc[i + B * k] = mk(0.0, 0.0);
  nk[0] = kept;
}

void davidson_block_orth(hipStream_t s, const cplx* G, cplx* Sc, int* nk) {
  hipLaunchKernelGGL(k_block_chol4, dim3(1), dim3(64), 0, s, G, Sc, nk);
  DC2_HIP(hipGetLastError());
}

void davidson_normalize(hipStream_t s, int64_t n, int ncol, cplx* T) {
  if (ncol <= 0) return;
  hipLaunchKernelGGL(k_col_normalize, dim3(ncol), dim3(256), 0, s, n, T);
  DC2_HIP(hipGetLastError());
}

void scale_rows(hipStream_t s, int64_t m, int64_t ncol, const double* f, const cplx* in, cplx* out) {
  if (m <= 0 || ncol <= 0) return;
  DC2_CHECK(ncol < 65536, 1, "scale_rows: too many columns for one grid dimension");
  hipLaunchKernelGGL(k_scale_rows, dim3(cdiv(m, 256), (unsigned)ncol), dim3(256), 0, s, m, ncol, f, in, out);
  DC2_HIP(hipGetLastError());
}

}  // namespace dc2
