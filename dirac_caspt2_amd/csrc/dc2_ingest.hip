// Integral ingest on the GPU: MDCINTNEW-equivalent records -> the dense blocks of the path.
//
// A record (i, j, k, l, v) holds one symmetry-unique two-electron integral (ij|kl) with
// 1-based CAS-order spinor labels, as create_newmdcint.F90:173-176 writes them.  Its 16
// images follow module_2integrals.F90:395-434: with t(p) the Kramers partner (p+1 for odd
// p, p-1 for even p), s_ij = +1 when i+j is even and -1 otherwise,
//   (ij|kl) = v, (t(j)t(i)|kl) = s_ij v, (ij|t(l)t(k)) = s_kl v, (t(j)t(i)|t(l)t(k)) = s_ij s_kl v,
//   (ji|lk) = conj v and its three partners, (kl|ij) = v and its three partners,
//   (lk|ji) = conj v and its three partners.
// Every image is written into each block whose space pattern it matches (OOOO, F1, F2,
// A1 ... H), which replaces the reference's dense fill plus the subspace files written
// by divide_2_elec_integral_into_subspaces.F90.  Where two records reach the same element
// the later record wins (then the later image), as in a sequential fill: a first pass
// takes the maximum key (record * 16 + image) per element with an atomic max, a second
// pass lets only the winner write.
//
// Before it, mdcint_convert turns DIRAC's unique integral set (MDCINT records) into those
// records on the device, as create_newmdcint.F90 does on the host: one thread per candidate
// entry, a count / scan / write pass keeping the reference's write order.
#include "dc2_kernels.hpp"

namespace dc2 {
namespace {

struct IngestBlock {
  cplx* data;
  unsigned long long* owner;
  int space[4];      // 0 inactive, 1 active, 2 secondary, 3 inactive+active ("o")
  int n[4];
};

constexpr int MAX_BLOCKS = 16;

struct IngestParams {
  int ni, na, ns, nblocks;
  IngestBlock b[MAX_BLOCKS];
};

__device__ inline int space_of(int p0, int ni, int na) {   // p0: 0-based spinor
  return p0 < ni ? 0 : (p0 < ni + na ? 1 : 2);
}

__device__ inline bool match(int want, int sp) { return want == 3 ? sp <= 1 : want == sp; }

__device__ inline int local_index(int want, int p0, int ni, int na) {
  return (want == 0 || want == 3) ? p0 : (want == 1 ? p0 - ni : p0 - ni - na);
}

__device__ inline int partner(int p) { return (p & 1) ? p + 1 : p - 1; }   // 1-based labels

// image g (0..15) of record (i, j, k, l, v): 1-based labels and value
__device__ inline void image(int g, int i, int j, int k, int l, cplx v, int* q, cplx& x) {
  const double sij = ((i + j) & 1) ? -1.0 : 1.0, skl = ((k + l) & 1) ? -1.0 : 1.0;
  const int it = partner(i), jt = partner(j), kt = partner(k), lt = partner(l);
  const cplx vc = cconj(v);
  switch (g) {
    case 0: q[0] = i; q[1] = j; q[2] = k; q[3] = l; x = v; break;
    case 1: q[0] = jt; q[1] = it; q[2] = k; q[3] = l; x = cscale(v, sij); break;
    case 2: q[0] = i; q[1] = j; q[2] = lt; q[3] = kt; x = cscale(v, skl); break;
    case 3: q[0] = jt; q[1] = it; q[2] = lt; q[3] = kt; x = cscale(v, sij * skl); break;
    case 4: q[0] = j; q[1] = i; q[2] = l; q[3] = k; x = vc; break;
    case 5: q[0] = it; q[1] = jt; q[2] = l; q[3] = k; x = cscale(vc, sij); break;
    case 6: q[0] = j; q[1] = i; q[2] = kt; q[3] = lt; x = cscale(vc, skl); break;
    case 7: q[0] = it; q[1] = jt; q[2] = kt; q[3] = lt; x = cscale(vc, sij * skl); break;
    case 8: q[0] = k; q[1] = l; q[2] = i; q[3] = j; x = v; break;
    case 9: q[0] = lt; q[1] = kt; q[2] = i; q[3] = j; x = cscale(v, skl); break;
    case 10: q[0] = k; q[1] = l; q[2] = jt; q[3] = it; x = cscale(v, sij); break;
    case 11: q[0] = lt; q[1] = kt; q[2] = jt; q[3] = it; x = cscale(v, sij * skl); break;
    case 12: q[0] = l; q[1] = k; q[2] = j; q[3] = i; x = vc; break;
    case 13: q[0] = kt; q[1] = lt; q[2] = j; q[3] = i; x = cscale(vc, skl); break;
    case 14: q[0] = l; q[1] = k; q[2] = it; q[3] = jt; x = cscale(vc, sij); break;
    default: q[0] = kt; q[1] = lt; q[2] = it; q[3] = jt; x = cscale(vc, sij * skl); break;
  }
}

template <int PASS>
__global__ void k_ingest(IngestParams P, int64_t nrec, const int* __restrict__ ijkl, const double* __restrict__ val) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrec * 16) return;
  const int64_t r = t >> 4;
  const int g = (int)(t & 15);
  const int nmo = P.ni + P.na + P.ns;
  const int i = ijkl[4 * r], j = ijkl[4 * r + 1], k = ijkl[4 * r + 2], l = ijkl[4 * r + 3];
  if (i < 1 || j < 1 || k < 1 || l < 1 || i > nmo || j > nmo || k > nmo || l > nmo) return;
  int q[4];
  cplx x;
  image(g, i, j, k, l, mk(val[2 * r], val[2 * r + 1]), q, x);
  int sp[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if (q[a] < 1 || q[a] > nmo) return;   // a partner label beyond the spinor set
    q[a] -= 1;
    sp[a] = space_of(q[a], P.ni, P.na);
  }
  const unsigned long long key = (unsigned long long)t;   // record * 16 + image
  for (int b = 0; b < P.nblocks; ++b) {
    const IngestBlock& B = P.b[b];
    if (!(match(B.space[0], sp[0]) && match(B.space[1], sp[1]) && match(B.space[2], sp[2]) &&
          match(B.space[3], sp[3])))
      continue;
    const int64_t e = (int64_t)local_index(B.space[0], q[0], P.ni, P.na) +
                      (int64_t)B.n[0] * (local_index(B.space[1], q[1], P.ni, P.na) +
                                         (int64_t)B.n[1] * (local_index(B.space[2], q[2], P.ni, P.na) +
                                                            (int64_t)B.n[2] * local_index(B.space[3], q[3], P.ni, P.na)));
    if (PASS == 0) atomicMax(&B.owner[e], key + 1);
    else if (B.owner[e] == key + 1) B.data[e] = x;
  }
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// ------------------------------------------------------------------ MDCINT -> records
// Candidate c of the conversion is entry (record r, image n, loop index i) in the order
// create_newmdcint.F90:86-177 visits them: records in file order; n = 0 the record as read,
// n = 1 its Kramers image (every label negated, re -> -re s, im -> im s with s the sign of
// ikr jkr indk indl); i over the nz entries and, when the SCHEME amplifies (scheme != 6 and
// |ikr| != |jkr|), over the nz particle-exchanged entries after them ((ij|kl) -> (ji|lk),
// every label negated for the types with one barred label on each side, else im -> -im).
// coff[r] = first candidate of record r (coff[nrec] = all).
struct MdcintDev {
  int64_t nrec, nkr;
  int scheme;
  const int64_t *ikr, *jkr, *nz, *eoff, *coff;
  const int *indk, *indl;
  const double *rr, *ri;   // ri null: real integrals
  const int* cas;           // cas[nkr + m]: CAS index of Kramers label m (d2c after kr)
};

struct MdOut {
  int q[4];
  double re, im;
  bool keep;
};

__device__ inline int64_t kramers_label(int64_t x) {   // 1 -1 2 -2 ... from CAS index 1 2 3 4 ...
  return (x & 1) ? x / 2 + 1 : -(x / 2);
}

// the record of candidate c: a binary search (lane 0 of the wave, for the wave's first
// candidate) then a forward walk over the records that end inside the wave (usually none:
// a record has 2 nz or 4 nz candidates)
__device__ inline int64_t mdcint_record(const MdcintDev& m, int64_t c, int64_t c0) {
  int64_t r0 = 0;
  if ((threadIdx.x & 63) == 0) {
    int64_t lo = 0, hi = m.nrec;   // last record with coff <= c0
    while (hi - lo > 1) {
      const int64_t md = (lo + hi) >> 1;
      if (m.coff[md] <= c0) lo = md;
      else hi = md;
    }
    r0 = lo;
  }
  int64_t r = __shfl(r0, 0, 64);
  while (r + 1 < m.nrec && m.coff[r + 1] <= c) ++r;
  return r;
}

__device__ MdOut mdcint_candidate(const MdcintDev& m, int64_t c, int64_t c0) {
  const int64_t r = mdcint_record(m, c, c0), nz = m.nz[r];
  int64_t ci = m.ikr[r], cj = m.jkr[r];
  const bool amplify = m.scheme != 6 && llabs(ci) != llabs(cj);
  const int64_t half = (amplify ? 2 : 1) * nz;
  const int64_t loc = c - m.coff[r];
  const int64_t n = loc / half, i = loc % half;
  const int64_t e = m.eoff[r] + i % nz;
  int64_t ck = m.indk[e], cl = m.indl[e];
  double re = m.rr[e], im = m.ri ? m.ri[e] : 0.0;
  if (n == 1) {
    const int neg = (ci < 0) + (cj < 0) + (ck < 0) + (cl < 0);
    const double sg = (neg & 1) ? -1.0 : 1.0;
    ci = -ci; cj = -cj; ck = -ck; cl = -cl;
    re = -re * sg;
    im = im * sg;
  }
  const double mag_re = fabs(re), mag_im = fabs(im);   // the cutoff reads the entry as stored
  if (i >= nz) {
    int64_t t = ci; ci = cj; cj = t;
    t = ck; ck = cl; cl = t;
    if ((ci < 0) + (cj < 0) == 1 && (ck < 0) + (cl < 0) == 1) {
      ci = -ci; cj = -cj; ck = -ck; cl = -cl;
    } else {
      im = -im;
    }
  }
  MdOut o;
  const int64_t lab[4] = {ci, cj, ck, cl};
  int64_t kl[4], ab[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t x = m.cas[m.nkr + lab[a]];   // CAS index
    kl[a] = kramers_label(x);
    ab[a] = llabs(kl[a]);
    o.q[a] = (int)((x & 1) ? x + 1 : x - 1);        // the Kramers partner, as written to MDCINTNEW
  }
  const double cutoff = 0.25e-12;
  bool ok = m.ri ? !(mag_re <= cutoff && mag_im <= cutoff) : !(mag_re <= cutoff);
  if (ok) {
    const int64_t a = ab[0], b = ab[1], cc = ab[2], d = ab[3];
    if (kl[0] > 0 && kl[1] > 0 && kl[2] > 0 && kl[3] > 0)
      ok = (a <= b && cc <= d && (a < cc || (a == cc && b <= d))) || (a <= b && d <= cc && (a < d || (a == d && b <= cc)));
    else if (kl[0] > 0 && kl[1] < 0 && kl[2] > 0 && kl[3] < 0)
      ok = a <= b && cc <= d && (a < cc || (a == cc && b <= d));
    else if (kl[0] > 0 && kl[1] < 0 && kl[2] < 0 && kl[3] > 0)
      ok = a <= b && cc <= d && (a < cc || (a == cc && b <= d));
    else if (kl[0] > 0 && kl[1] < 0 && kl[2] < 0 && kl[3] < 0)
      ok = a <= b;
    else
      ok = false;
  }
  o.re = re;
  o.im = -im;
  o.keep = ok;
  return o;
}

constexpr int MD_BLOCK = 256;

__global__ __launch_bounds__(MD_BLOCK) void k_md_count(MdcintDev m, int64_t ncand, int64_t* __restrict__ bcount) {
  const int64_t c = (int64_t)blockIdx.x * MD_BLOCK + threadIdx.x;
  const int64_t c0 = c - (threadIdx.x & 63);   // the wave's first candidate (< ncand)
  bool keep = false;
  if (c0 < ncand) {   // uniform over the wave: every lane takes part in the search's shuffle
    const MdOut o = mdcint_candidate(m, c < ncand ? c : c0, c0);
    keep = c < ncand && o.keep;
  }
  const int k = __syncthreads_count(keep);
  if (threadIdx.x == 0) bcount[blockIdx.x] = k;
}

// exclusive scan of nb block counts in one workgroup of 1024 threads; boff[nb] = the total
__global__ __launch_bounds__(1024) void k_md_scan(int64_t nb, const int64_t* __restrict__ bcount,
                                                  int64_t* __restrict__ boff) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (nb + 1023) / 1024, b0 = t * per, b1 = min(nb, b0 + per);
  int64_t sum = 0;
  for (int64_t b = b0; b < b1; ++b) sum += bcount[b];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {   // inclusive Hillis-Steele scan
    const int64_t v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t run = part[t] - sum;
  for (int64_t b = b0; b < b1; ++b) {
    boff[b] = run;
    run += bcount[b];
  }
  if (t == 1023) boff[nb] = part[1023];
}

__global__ __launch_bounds__(MD_BLOCK) void k_md_write(MdcintDev m, int64_t ncand, const int64_t* __restrict__ boff,
                                                       int* __restrict__ ijkl, double* __restrict__ val) {
  __shared__ int wsum[MD_BLOCK / 64];
  const int64_t c = (int64_t)blockIdx.x * MD_BLOCK + threadIdx.x;
  const int64_t c0 = c - (threadIdx.x & 63);
  MdOut o;
  o.keep = false;
  if (c0 < ncand) {
    o = mdcint_candidate(m, c < ncand ? c : c0, c0);
    if (c >= ncand) o.keep = false;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long mask = __ballot(o.keep);
  const int below = __popcll(mask & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[wave] = __popcll(mask);
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += wsum[w];
  if (o.keep) {
    const int64_t p = boff[blockIdx.x] + before + below;
#pragma unroll
    for (int a = 0; a < 4; ++a) ijkl[4 * p + a] = o.q[a];
    val[2 * p] = o.re;
    val[2 * p + 1] = o.im;
  }
}

}  // namespace

int64_t mdcint_convert_scratch(int64_t ncand) { return 8 * (2 * cdiv(ncand, MD_BLOCK) + 1); }

void mdcint_convert(hipStream_t s, int64_t nrec, const int64_t* ikr, const int64_t* jkr, const int64_t* nz,
                    const int64_t* eoff, const int64_t* coff, const int* indk, const int* indl,
                    const double* rr, const double* ri, int64_t nkr, const int* cas, int scheme, int64_t ncand,
                    int64_t* scratch, int* ijkl, double* val) {
  const int64_t nb = cdiv(ncand, MD_BLOCK);
  DC2_CHECK(nb < (1ll << 31), 1, "mdcint: too many entries for one launch");
  MdcintDev m{nrec, nkr, scheme, ikr, jkr, nz, eoff, coff, indk, indl, rr, ri, cas};
  int64_t* bcount = scratch;
  int64_t* boff = scratch + nb;
  if (nb > 0) {
    hipLaunchKernelGGL(k_md_count, dim3(nb), dim3(MD_BLOCK), 0, s, m, ncand, bcount);
    hipLaunchKernelGGL(k_md_scan, dim3(1), dim3(1024), 0, s, nb, bcount, boff);
    hipLaunchKernelGGL(k_md_write, dim3(nb), dim3(MD_BLOCK), 0, s, m, ncand, boff, ijkl, val);
  } else {
    hipLaunchKernelGGL(k_md_scan, dim3(1), dim3(1024), 0, s, nb, bcount, boff);
  }
  DC2_HIP(hipGetLastError());
}

void ingest_records(hipStream_t s, int ni, int na, int ns, int nblocks, const int* const* spaces, const int* const* dims,
                    cplx* const* data, unsigned long long* const* owner, const int64_t* counts, int64_t nrec,
                    const int* ijkl, const double* val) {
  DC2_CHECK(nblocks >= 1 && nblocks <= MAX_BLOCKS, 1, "ingest: bad block count");
  DC2_CHECK(nrec >= 0 && cdiv(nrec * 16, 256) < (1ll << 31), 1, "ingest: too many records for one launch");
  IngestParams P;
  P.ni = ni; P.na = na; P.ns = ns; P.nblocks = nblocks;
  for (int b = 0; b < nblocks; ++b) {
    P.b[b].data = data[b];
    P.b[b].owner = owner[b];
    for (int a = 0; a < 4; ++a) { P.b[b].space[a] = spaces[b][a]; P.b[b].n[a] = dims[b][a]; }
    fill_zero(s, data[b], counts[b]);
    if (counts[b] > 0) hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(counts[b], 256)), dim3(256), 0, s, owner[b], counts[b]);
  }
  if (nrec > 0) {
    const int64_t nt = nrec * 16;
    hipLaunchKernelGGL(k_ingest<0>, dim3(cdiv(nt, 256)), dim3(256), 0, s, P, nrec, ijkl, val);
    hipLaunchKernelGGL(k_ingest<1>, dim3(cdiv(nt, 256)), dim3(256), 0, s, P, nrec, ijkl, val);
  }
  DC2_HIP(hipGetLastError());
}

}  // namespace dc2
