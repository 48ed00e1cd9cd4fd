// Vectorised block tiles (included by kernels.hip inside namespace tpamr).
//
// A thread owns a 4-cell segment of one block row and moves it with 16-byte
// loads and stores; a 256-thread workgroup holds BPW = 256 / (NB/4) blocks
// (one wavefront per 16x16 block). Phase A issues every global load of the
// thread (segment, conductances, halo or external slot, its irregular row,
// the caller's operands); one barrier; phase B lets one lane per irregular
// row sum its row from LDS; phase C evaluates the 5-point stencil of the four
// cells from registers, 16-byte LDS reads of the rows above and below, and
// lane shuffles for the west and east neighbours.
#pragma once

template <typename T>
struct Vec4 {
  T v[4];
};

template <typename T>
__device__ inline Vec4<T> ld4(const T* __restrict__ p) {
  Vec4<T> r;
  if constexpr (sizeof(T) == 4) {
    const float4 q = *reinterpret_cast<const float4*>(p);
    r.v[0] = q.x;
    r.v[1] = q.y;
    r.v[2] = q.z;
    r.v[3] = q.w;
  } else {
    const double2 a = *reinterpret_cast<const double2*>(p);
    const double2 c = *reinterpret_cast<const double2*>(p + 2);
    r.v[0] = a.x;
    r.v[1] = a.y;
    r.v[2] = c.x;
    r.v[3] = c.y;
  }
  return r;
}

template <typename T>
__device__ inline void st4(T* __restrict__ p, const Vec4<T>& r) {
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<float4*>(p) = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
  } else {
    *reinterpret_cast<double2*>(p) = make_double2(r.v[0], r.v[1]);
    *reinterpret_cast<double2*>(p + 2) = make_double2(r.v[2], r.v[3]);
  }
}

template <int BX, int BY>
struct Geo {
  static constexpr int NB = BX * BY;
  static constexpr int TPR = BX / 4;                 // threads per block row
  static constexpr int TPB = NB / 4;                 // threads per block
  static constexpr int BPW = TPB >= 256 ? 1 : 256 / TPB;  // blocks per workgroup
  static constexpr int WG = BPW * TPB;
  static constexpr int HALO = 2 * (BX + BY);
  static constexpr int SEG = TPB < 64 ? TPB : 64;    // lanes of one block within a wavefront
  // LDS value slots of a block tile: the interior cells (row-major, like the
  // cells in HBM), the halo strips W and E (one value per block row), S and N
  // (one per block column), then the external cells of the irregular rows.
  // The irregular rows' column references (built on the host) use this
  // numbering.
  static constexpr int HW = NB, HE = NB + BY, HS = NB + 2 * BY, HN = NB + 2 * BY + BX, EXT = NB + HALO;
};

// Irregular rows: entries in groups of four per row (ELL_G groups at most),
// each group one 16-byte load of values and one 8-byte load of four 16-bit
// LDS slot references; padding entries have value 0 and reference the row's
// own cell.
constexpr int ELL_G = ELL_MAX / 4;

// LDS slot of one block: tile values (interior, halo strips, `ext` external
// cells), the south conductances of the block (north exchange), and the
// irregular rows' products, diagonals and flags. Slots are carved from the
// dynamic LDS of the launch (bytes(ext) each, BPW slots per workgroup); every
// array starts 16-byte aligned.
template <typename T, int BX, int BY>
struct Slot4 {
  T* v;
  T* cs;
  T* ax;
  T* d;
  unsigned* flag4;  // one byte per cell, four cells per word
  static __host__ __device__ inline int bytes(int ext) {
    using G = Geo<BX, BY>;
    const int vals = G::EXT + ((ext + 3) & ~3) + 3 * G::NB;
    return ((vals * int(sizeof(T)) + G::NB + 15) / 16) * 16;
  }
  __device__ inline Slot4(unsigned char* lds, int slot, int ext) {
    using G = Geo<BX, BY>;
    unsigned char* base = lds + slot * bytes(ext);
    v = reinterpret_cast<T*>(base);
    cs = v + G::EXT + ((ext + 3) & ~3);
    ax = cs + G::NB;
    d = ax + G::NB;
    flag4 = reinterpret_cast<unsigned*>(d + G::NB);
  }
};

// plain array fetches
template <typename T>
struct Arr4 {
  const T* __restrict__ v;
  __device__ inline Vec4<T> seg(int cell0) const { return ld4(v + cell0); }
  __device__ inline T one(int cell) const { return v[cell]; }
};

// the zero-start Jacobi iterate x0 = (w D^-1) b of a level, formed where it
// is read (the same products as the Krylov update that would store it)
template <typename T>
struct ZeroStart4 {
  const T* __restrict__ b;
  const T* __restrict__ dinv;
  T w;
  __device__ inline Vec4<T> seg(int cell0) const {
    const Vec4<T> bv = ld4(b + cell0), dv = ld4(dinv + cell0);
    Vec4<T> r;
#pragma unroll
    for (int m = 0; m < 4; ++m) r.v[m] = (w * dv.v[m]) * bv.v[m];
    return r;
  }
  __device__ inline T one(int cell) const { return (w * dinv[cell]) * b[cell]; }
};

// injected coarse values t = P_t x_c (identity block or one quadrant)
template <typename T, int BX, int BY>
struct Inject4 {
  const T* __restrict__ xc;
  const int* __restrict__ xfer;
  __device__ inline T one(int cell) const { return xc[owner_of<BX, BY>(cell, xfer)]; }
  __device__ inline Vec4<T> seg(int cell0) const {
    constexpr int NB = BX * BY;
    const int blk = cell0 / NB, loc = cell0 - blk * NB;
    const int cb = xfer[2 * blk], mode = xfer[2 * blk + 1];
    if (mode < 0) return ld4(xc + cb * NB + loc);
    const int i = loc % BX, j = loc / BX;
    const int ci = (mode & 1) * (BX / 2) + (i >> 1), cj = (mode >> 1) * (BY / 2) + (j >> 1);
    const T a = xc[cb * NB + cj * BX + ci], c = xc[cb * NB + cj * BX + ci + 1];
    Vec4<T> r;
    r.v[0] = a;
    r.v[1] = a;
    r.v[2] = c;
    r.v[3] = c;
    return r;
  }
};

struct Pre0 {
  __device__ inline void operator()() const {}
};

// Per-thread placement: thread t (0..255) of tile group `wg` (BPW blocks),
// using the LDS slots of quarter `quarter` (a tail launch runs `span` tile
// groups at once, one per 256 threads: groups wg - quarter ... + span - 1).
// The irregular-row phase runs when any block of those groups has irregular
// rows, so its barrier is reached by the whole workgroup.
template <int BX, int BY>
struct Place {
  int wg, b, lt, j, q, i0, cell0, slot_, blk_lo, blk_hi;
  bool valid;
  __device__ inline Place(int wg_, int nblocks, int t, int quarter, int span = 1) : wg(wg_) {
    using G = Geo<BX, BY>;
    blk_lo = min((wg - quarter) * G::BPW, nblocks);
    blk_hi = min(blk_lo + span * G::BPW, nblocks);
    const int k = t / G::TPB;
    lt = t - k * G::TPB;
    b = wg * G::BPW + k;
    // a block fills whole wavefronts when TPB >= 64: keep its index scalar
    if constexpr (G::TPB >= 64) b = __builtin_amdgcn_readfirstlane(b);
    valid = b < nblocks;
    j = lt / G::TPR;
    q = lt - j * G::TPR;
    i0 = 4 * q;
    cell0 = b * G::NB + j * BX + i0;
    slot_ = quarter * G::BPW + k;
  }
  __device__ inline int slot() const { return slot_; }
};

// Evaluate (A v) and D of the thread's four cells (or, with ABS, the
// factorized round-off bound sum_j B_cj |v_j|). `vc` returns the cells' own
// values (absolute values with ABS). The lane that owns an irregular row
// writes its product, diagonal and flag to LDS; the owners of the cells read
// them after the second barrier.
template <typename T, int BX, int BY, bool ABS, class F, class Pre>
__device__ inline void tile4(unsigned char* lds, const LevelView<T>& L, const Place<BX, BY>& P, F fetch, Pre pre,
                             Vec4<T>& vc, Vec4<T>& ax, Vec4<T>& d) {
  using G = Geo<BX, BY>;
  constexpr int NB = G::NB, TPB = G::TPB, TPR = G::TPR;
  const Slot4<T, BX, BY> S(lds, P.slot(), L.ext_cap);
  const int b = P.b, lt = P.lt, j = P.j, q = P.q, i0 = P.i0;
  const int seg = 4 * lt;  // the segment's first interior slot (j * BX + i0)
  // rows of the workgroup's blocks: blk_irr is a prefix sum over blocks
  const bool any = L.blk_irr[P.blk_hi] > L.blk_irr[P.blk_lo];
  const int ng = (L.ell_w + 3) >> 2;  // entry groups per row
  const T* __restrict__ vals = ABS ? L.ell_bound : L.ell_val;
  Vec4<T> own{}, cw4{}, cs4{}, cn4{};
  T rdg = 0, rbd = 0;
  T ceb = 0;
  int nr = 0, r0 = 0;
  Vec4<T> rv[ELL_G];
  uint2 rr[ELL_G];
  int rloc = 0, rlen = 0;
  unsigned bits = 0;
  if (P.valid) {
    own = fetch.seg(P.cell0);
    cw4 = ld4(L.cw + P.cell0);
    cs4 = ld4(L.cs + P.cell0);
    if (q == TPR - 1) ceb = L.ce[b * BY + j];
    if (j == BY - 1) cn4 = ld4(L.cn + b * BX + i0);
    pre();  // the kernel's own loads (no LDS)
    // halo values (W, E, S, N) of same-level neighbour blocks: halo index h
    // is LDS slot NB + h
    auto halo_value = [&](int h) {
      int side, src;
      if (h < BY) {
        side = 0; src = h * BX + (BX - 1);
      } else if (h < 2 * BY) {
        side = 1; src = (h - BY) * BX;
      } else if (h < 2 * BY + BX) {
        side = 2; src = (BY - 1) * BX + (h - 2 * BY);
      } else {
        side = 3; src = h - 2 * BY - BX;
      }
      const int nb = L.nbr[4 * b + side];
      return nb >= 0 ? fetch.one(nb * NB + src) : T(0);
    };
    // the lane's first halo value and first external cell are loaded into
    // registers and written to LDS after every load of the tile is issued,
    // so the irregular rows' loads do not wait for the halo values
    const T hv0 = lt < G::HALO ? halo_value(lt) : T(0);
    T ev0 = 0;
    int e0 = 0, ne = 0;
    if (any) {
      // this lane's irregular row (rows in decreasing length order; group g
      // of row x of the block at (r0 * ng + g * nr + x) * 4). Its loads are
      // predicated on the block's group count, not on the row's own length,
      // so they wait on no other load than the block's row offsets; padding
      // entries of shorter rows are loaded, not used.
      r0 = L.blk_irr[b];
      nr = L.blk_irr[b + 1] - r0;
      int ngb = L.blk_ng[b];
      if constexpr (G::TPB >= 64) {
        // uniform per wavefront: group offsets become scalar
        r0 = __builtin_amdgcn_readfirstlane(r0);
        nr = __builtin_amdgcn_readfirstlane(nr);
        ngb = __builtin_amdgcn_readfirstlane(ngb);
      }
      unsigned info = 0;
      if (lt < nr) {
        info = L.ell_loc[r0 + lt];
        rdg = L.ell_diag[r0 + lt];
        if constexpr (ABS) rbd = L.ell_bdiag[r0 + lt];
        const int base = (r0 * ng + lt) * 4;
#pragma unroll
        for (int g = 0; g < ELL_G; ++g)
          if (g < ngb) {
            const int at = base + g * nr * 4;
            rv[g] = ld4(vals + at);
            rr[g] = *reinterpret_cast<const uint2*>(L.ell_ref + at);
          }
      }
      // external cells of the irregular rows
      e0 = L.blk_ext[b];
      ne = L.blk_ext[b + 1] - e0;
      if (lt < ne) ev0 = fetch.one(L.ext_col[e0 + lt]);
      rloc = int(info & 0xffffu);
      rlen = int(info >> 16);
    }
    if (lt < G::HALO) S.v[NB + lt] = ABS ? fabs(hv0) : hv0;
    for (int h = lt + TPB; h < G::HALO; h += TPB) {
      const T hv = halo_value(h);
      S.v[NB + h] = ABS ? fabs(hv) : hv;
    }
    if (any) {
      if (lt < ne) S.v[G::EXT + lt] = ABS ? fabs(ev0) : ev0;
      for (int e = lt + TPB; e < ne; e += TPB) {
        const T ev = fetch.one(L.ext_col[e0 + e]);
        S.v[G::EXT + e] = ABS ? fabs(ev) : ev;
      }
      S.flag4[lt] = 0u;  // this lane's four cells
    }
    if constexpr (ABS) {
#pragma unroll
      for (int m = 0; m < 4; ++m) own.v[m] = fabs(own.v[m]);
    }
    st4(S.v + seg, own);
    st4(S.cs + seg, cs4);
  }
  __syncthreads();
  if (any) {
    if (P.valid) {
      // The row's ELL part (the couplings of its non-regular faces) sums to
      // zero, like each face's contribution, so it is evaluated in flux form
      // sum_j a_ij (v_j - v_i) over its off-diagonal entries: the zero padding
      // drops out and rounding scales with the differences, as in the
      // reference's face-based product. The stencil adds the cell's regular
      // and open faces in phase C. ABS keeps sum_j |b_ij||v_j| with the
      // bound's diagonal part rbd |v_i|.
      if (lt < nr) {
        const T vi = ABS ? T(0) : S.v[rloc];
        T s = 0;
#pragma unroll
        for (int g = 0; g < ELL_G; ++g)
          if (g < ng && 4 * g < rlen) {
            s += rv[g].v[0] * (S.v[rr[g].x & 0xffffu] - vi);
            s += rv[g].v[1] * (S.v[rr[g].x >> 16] - vi);
            s += rv[g].v[2] * (S.v[rr[g].y & 0xffffu] - vi);
            s += rv[g].v[3] * (S.v[rr[g].y >> 16] - vi);
          }
        if constexpr (ABS) s += rbd * S.v[rloc];
        S.ax[rloc] = s;
        S.d[rloc] = rdg;
        reinterpret_cast<unsigned char*>(S.flag4)[rloc] = 1;
      }
      // blocks with more irregular rows than lanes
      for (int x = lt + TPB; x < nr; x += TPB) {
        const unsigned info = L.ell_loc[r0 + x];
        const int loc = int(info & 0xffffu), len = int(info >> 16);
        const T vi = ABS ? T(0) : S.v[loc];
        T s = 0;
        for (int g = 0; 4 * g < len; ++g) {
          const int at = (r0 * ng + x) * 4 + g * nr * 4;
          const Vec4<T> v = ld4(vals + at);
          const uint2 r = *reinterpret_cast<const uint2*>(L.ell_ref + at);
          s += v.v[0] * (S.v[r.x & 0xffffu] - vi);
          s += v.v[1] * (S.v[r.x >> 16] - vi);
          s += v.v[2] * (S.v[r.y & 0xffffu] - vi);
          s += v.v[3] * (S.v[r.y >> 16] - vi);
        }
        if constexpr (ABS) s += L.ell_bdiag[r0 + x] * S.v[loc];
        S.ax[loc] = s;
        S.d[loc] = L.ell_diag[r0 + x];
        reinterpret_cast<unsigned char*>(S.flag4)[loc] = 1;
      }
    }
    __syncthreads();
  }
  // west / east neighbours inside the block row come from the next lanes
  const T cw_next = __shfl_down(cw4.v[0], 1, G::SEG);
  const T v_prev = __shfl_up(own.v[3], 1, G::SEG);
  const T v_next = __shfl_down(own.v[0], 1, G::SEG);
  vc = own;
  if (!P.valid) return;
  const T vwest = q > 0 ? v_prev : S.v[G::HW + j];
  const T veast = q < TPR - 1 ? v_next : S.v[G::HE + j];
  const Vec4<T> vs4 = ld4(S.v + (j > 0 ? seg - BX : G::HS + i0));
  const Vec4<T> vn4 = ld4(S.v + (j < BY - 1 ? seg + BX : G::HN + i0));
  if (j < BY - 1) cn4 = ld4(S.cs + seg + BX);
  Vec4<T> axi{}, dgi{};
  if (any) {
    const unsigned f = S.flag4[lt];
    bits = (f & 1u) | ((f >> 7) & 2u) | ((f >> 14) & 4u) | ((f >> 21) & 8u);
    if (bits) {
      axi = ld4(S.ax + seg);
      dgi = ld4(S.d + seg);
    }
  }
  // stencil of every cell (an irregular cell's regular and open faces), plus
  // the ELL part of the irregular ones (keeps the results in registers)
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const T vw = m > 0 ? own.v[m - 1] : vwest;
    const T ve = m < 3 ? own.v[m + 1] : veast;
    const T vs = vs4.v[m], vn = vn4.v[m];
    const T cw = cw4.v[m];
    const T ce = m < 3 ? cw4.v[m + 1] : (q < TPR - 1 ? cw_next : ceb);
    const T cs = cs4.v[m];
    const T cn = cn4.v[m];
    const T dd = (cw + ce) + (cs + cn);
    const T vi = own.v[m];
    const T st = ABS ? dd * vi + ((cw * vw + ce * ve) + (cs * vs + cn * vn))
                     : (cw * (vi - vw) + ce * (vi - ve)) + (cs * (vi - vs) + cn * (vi - vn));
    const bool irr = (bits >> m) & 1u;
    ax.v[m] = irr ? st + axi.v[m] : st;
    d.v[m] = irr ? dgi.v[m] : dd;
  }
}

// Sum of `v` over the lanes of one block; valid in lane lt == 0.
template <typename T, int BX, int BY>
__device__ inline T block4_sum(T v, T* scratch) {
  using G = Geo<BX, BY>;
  for (int o = G::SEG / 2; o > 0; o >>= 1) v += __shfl_down(v, o, G::SEG);
  if constexpr (G::TPB <= 64) {
    return v;
  } else {
    constexpr int NW = G::TPB / 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) scratch[wv] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0)
      for (int k = 0; k < NW; ++k) s += scratch[k];
    __syncthreads();
    return s;
  }
}

template <typename T, int BX, int BY>
__device__ inline T block4_max(T v, T* scratch) {
  using G = Geo<BX, BY>;
  for (int o = G::SEG / 2; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, G::SEG));
  if constexpr (G::TPB <= 64) {
    return v;
  } else {
    constexpr int NW = G::TPB / 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) scratch[wv] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0)
      for (int k = 0; k < NW; ++k) s = fmax(s, scratch[k]);
    __syncthreads();
    return s;
  }
}

#define TILE4_SHARED(T, BX, BY) extern __shared__ __align__(16) unsigned char S_[]

// Minimum waves per SIMD the tile kernels are compiled for: FP32 tiles fit
// 8 (64 VGPRs; their LDS slots allow 8 workgroups per CU), FP64 tiles keep
// the compiler's choice (their LDS slots allow 4).
#define TILE_WAVES(T) (sizeof(T) == 4 ? 8 : 1)

// Damped Jacobi sweep x_out = x_in + w D^-1 (b - A x_in); STATS: per-block
// partial sums of V z, r z and r for the Krylov update (r of type TR: the
// FP64 Krylov residual or the FP32 inner residual).
template <typename T, int BX, int BY, bool STATS, typename TR, class F = Arr4<T>>
__device__ inline void sweep_op(unsigned char* lds, const LevelView<T>& L, const Place<BX, BY>& P, F x_in,
                                T* __restrict__ x_out, const T* __restrict__ rhs, double omega,
                                const TR* __restrict__ rz, double* __restrict__ part) {
  Vec4<T> bv{}, vc, ax, d;
  Vec4<double> rd{};
  tile4<T, BX, BY, false>(lds, L, P, x_in, [&]() {
    bv = ld4(rhs + P.cell0);
    if constexpr (STATS) {
      const Vec4<TR> q = ld4(rz + P.cell0);
#pragma unroll
      for (int m = 0; m < 4; ++m) rd.v[m] = double(q.v[m]);
    }
  }, vc, ax, d);
  Vec4<T> z;
  double sz = 0, srz = 0, sr = 0;
  if (P.valid) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      // a cell without conducting faces (zero diagonal, its own component)
      // keeps its value, as with the reference's zero inverse diagonal
      const T step = T(omega) * (bv.v[m] - ax.v[m]) / d.v[m];
      z.v[m] = d.v[m] > T(0) ? vc.v[m] + step : vc.v[m];
      if constexpr (STATS) {
        sz += double(z.v[m]);
        srz += rd.v[m] * double(z.v[m]);
        sr += rd.v[m];
      }
    }
    st4(x_out + P.cell0, z);
  }
  if constexpr (STATS) {
    __shared__ double red[4];
    sz = block4_sum<double, BX, BY>(sz, red);
    srz = block4_sum<double, BX, BY>(srz, red);
    sr = block4_sum<double, BX, BY>(sr, red);
    if (P.valid && P.lt == 0) {
      const int nb = L.nblocks;
      part[P.b] = sz * L.volume[P.b];
      part[nb + P.b] = srz;
      part[2 * nb + P.b] = sr;
    }
  }
}

// u = w_P D^-1 (b - A x); a cell without conducting faces (zero diagonal)
// keeps its residual b - A x in u instead, which its neighbours' products do
// not see (their couplings to it are zero)
template <typename T, int BX, int BY>
__device__ inline void residual_op(unsigned char* lds, const LevelView<T>& L, const Place<BX, BY>& P,
                                   const T* __restrict__ x, const T* __restrict__ rhs, T* __restrict__ u,
                                   double omega_p) {
  Vec4<T> bv{}, vc, ax, d;
  tile4<T, BX, BY, false>(lds, L, P, Arr4<T>{x}, [&]() { bv = ld4(rhs + P.cell0); }, vc, ax, d);
  if (!P.valid) return;
  Vec4<T> o;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const T r = bv.v[m] - ax.v[m];
    o.v[m] = d.v[m] > T(0) ? T(omega_p) * r / d.v[m] : r;
  }
  st4(u + P.cell0, o);
}

// Smoothed restriction b_c = P_t' (r - S u) with r = D u / w_P (the residual
// itself for a zero diagonal, whose smoothed transfer row is the tentative
// one), and the zero-start coarse sweep x_c = w_c D_c^-1 b_c.
template <typename T, int BX, int BY>
__device__ inline void restrict_op(unsigned char* lds, const LevelView<T>& L, const Place<BX, BY>& P,
                                   const T* __restrict__ u, double omega_p, T* __restrict__ bc, T* __restrict__ xc,
                                   const T* __restrict__ dinv_c, double omega_c, int zero_start) {
  using G = Geo<BX, BY>;
  Vec4<T> vc, ax, d;
  tile4<T, BX, BY, false>(lds, L, P, Arr4<T>{u}, Pre0{}, vc, ax, d);
  Vec4<T> w;
#pragma unroll
  for (int m = 0; m < 4; ++m)
    w.v[m] = P.valid ? (d.v[m] > T(0) ? d.v[m] * vc.v[m] / T(omega_p) - ax.v[m] : vc.v[m]) : T(0);
  const int cb = P.valid ? L.xfer[2 * P.b] : 0, mode = P.valid ? L.xfer[2 * P.b + 1] : -1;
  // pair sums across the row segment, then across row pairs (same wavefront)
  const T h0 = w.v[0] + w.v[1], h1 = w.v[2] + w.v[3];
  const T n0 = __shfl_down(h0, G::TPR, G::SEG), n1 = __shfl_down(h1, G::TPR, G::SEG);
  if (!P.valid) return;
  if (mode < 0) {
    const int cc = cb * G::NB + P.j * BX + P.i0;
    st4(bc + cc, w);
    Vec4<T> x0;
    if (zero_start) {
      const Vec4<T> di = ld4(dinv_c + cc);
#pragma unroll
      for (int m = 0; m < 4; ++m) x0.v[m] = T(omega_c) * di.v[m] * w.v[m];
    } else {
#pragma unroll
      for (int m = 0; m < 4; ++m) x0.v[m] = T(0);
    }
    st4(xc + cc, x0);
    return;
  }
  if (P.j & 1) return;
  const T s0 = h0 + n0, s1 = h1 + n1;
  const int ci = (mode & 1) * (BX / 2) + (P.i0 >> 1), cj = (mode >> 1) * (BY / 2) + (P.j >> 1);
  const int cc = cb * G::NB + cj * BX + ci;
  bc[cc] = s0;
  bc[cc + 1] = s1;
  if (zero_start) {
    xc[cc] = T(omega_c) * dinv_c[cc] * s0;
    xc[cc + 1] = T(omega_c) * dinv_c[cc + 1] * s1;
  } else {
    xc[cc] = T(0);
    xc[cc + 1] = T(0);
  }
}

// x += t - w_P D^-1 S t with t = P_t x_c
template <typename T, int BX, int BY>
__device__ inline void prolong_op(unsigned char* lds, const LevelView<T>& L, const Place<BX, BY>& P, T* __restrict__ x,
                                  const T* __restrict__ xc, double omega_p) {
  Vec4<T> xv{}, vc, ax, d;
  tile4<T, BX, BY, false>(lds, L, P, Inject4<T, BX, BY>{xc, L.xfer}, [&]() { xv = ld4(x + P.cell0); }, vc, ax, d);
  if (!P.valid) return;
  Vec4<T> o;
#pragma unroll
  for (int m = 0; m < 4; ++m) o.v[m] = xv.v[m] + (d.v[m] > T(0) ? vc.v[m] - T(omega_p) * ax.v[m] / d.v[m] : vc.v[m]);
  st4(x + P.cell0, o);
}

// Tile group of a standalone tile launch (padding workgroups return).
#define TILE_GROUP(L)                                       \
  const int wg = xcd_chunk_group(blockIdx.x, (L).wg_chunk);  \
  if (wg * Geo<BX, BY>::BPW >= (L).nblocks) return;           \
  const Place<BX, BY> P(wg, (L).nblocks, threadIdx.x, 0)

// ZS: the sweep starts from the zero-start iterate x0 = (w0 D^-1) rhs,
// formed from rhs and the level's D^-1 (x_in unused): two damped Jacobi
// sweeps from zero in one pass.
template <typename T, int BX, int BY, bool FINE, bool STATS, typename TR = double, bool ZS = false>
__global__ void __launch_bounds__(256, TILE_WAVES(T))
    k_sweep(LevelView<T> L, const T* __restrict__ x_in, T* __restrict__ x_out, const T* __restrict__ rhs, double omega,
            const int* __restrict__ done, const TR* __restrict__ rz, double* __restrict__ part,
            const T* __restrict__ dinv0, double w0) {
  if (done && *done) return;
  TILE4_SHARED(T, BX, BY);
  TILE_GROUP(L);
  if constexpr (ZS)
    sweep_op<T, BX, BY, STATS, TR>(S_, L, P, ZeroStart4<T>{rhs, dinv0, T(w0)}, x_out, rhs, omega, rz, part);
  else
    sweep_op<T, BX, BY, STATS, TR>(S_, L, P, Arr4<T>{x_in}, x_out, rhs, omega, rz, part);
}

template <typename T, int BX, int BY, bool FINE>
__global__ void __launch_bounds__(256, TILE_WAVES(T))
    k_residual(LevelView<T> L, const T* __restrict__ x, const T* __restrict__ rhs, T* __restrict__ u, double omega_p,
               const int* __restrict__ done) {
  if (done && *done) return;
  TILE4_SHARED(T, BX, BY);
  TILE_GROUP(L);
  residual_op<T, BX, BY>(S_, L, P, x, rhs, u, omega_p);
}

template <typename T, int BX, int BY, bool FINE>
__global__ void __launch_bounds__(256, TILE_WAVES(T))
    k_restrict(LevelView<T> L, const T* __restrict__ u, double omega_p, T* __restrict__ bc, T* __restrict__ xc,
               const T* __restrict__ dinv_c, double omega_c, int zero_start, const int* __restrict__ done) {
  if (done && *done) return;
  TILE4_SHARED(T, BX, BY);
  TILE_GROUP(L);
  restrict_op<T, BX, BY>(S_, L, P, u, omega_p, bc, xc, dinv_c, omega_c, zero_start);
}

template <typename T, int BX, int BY, bool FINE>
__global__ void __launch_bounds__(256, TILE_WAVES(T))
    k_prolong(LevelView<T> L, T* __restrict__ x, const T* __restrict__ xc, double omega_p,
              const int* __restrict__ done) {
  if (done && *done) return;
  TILE4_SHARED(T, BX, BY);
  TILE_GROUP(L);
  prolong_op<T, BX, BY>(S_, L, P, x, xc, omega_p);
}

// Coarsest level: x = inv * b with the dense inverse of the augmented
// (volume-constrained) operator, column-major, evaluated in FP64 by the
// whole workgroup: 64 rows at a time, blockDim / 64 column groups summed
// through LDS.
template <typename T>
__device__ inline void coarse_op(int n, const double* __restrict__ inv, const T* __restrict__ b, T* __restrict__ x) {
  __shared__ double part[16][65];
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6, ng = blockDim.x >> 6;
  for (int r0 = 0; r0 < n; r0 += 64) {
    const int r = r0 + lane;
    double s0 = 0, s1 = 0;
    if (r < n) {
      int c = g;
      for (; c + ng < n; c += 2 * ng) {
        s0 += inv[size_t(c) * n + r] * double(b[c]);
        s1 += inv[size_t(c + ng) * n + r] * double(b[c + ng]);
      }
      if (c < n) s0 += inv[size_t(c) * n + r] * double(b[c]);
    }
    part[g][lane] = s0 + s1;
    __syncthreads();
    if (g == 0 && r < n) {
      double s = 0;
      for (int q = 0; q < ng; ++q) s += part[q][lane];
      x[r] = T(s);
    }
    __syncthreads();
  }
}

// The V-cycle below a level small enough for one workgroup: levels
// 0..n-1 of `A` (the last one the coarsest), tile passes of up to
// blockDim / 256 tile groups at once separated by workgroup barriers. On
// entry b[0] and the zero-start x[0] come from the restriction of the next
// finer level; on exit level 0's result is in x[0] or xs[0] by the parity
// of its (pre - 1) + post sweeps, as in the launch sequence. Phase 1 stops
// once the coarsest level's rhs is restricted, phase 2 starts from its
// solution (the parities after the way down are fixed by `pre`).
template <typename T, int BX, int BY>
__global__ void __launch_bounds__(1024) k_vtail(TailArgs<T> A, const int* __restrict__ done) {
  if (done && *done) return;
  TILE4_SHARED(T, BX, BY);
  using G = Geo<BX, BY>;
  const int nq = blockDim.x >> 8, quarter = threadIdx.x >> 8, t = threadIdx.x & 255;
  // bit l: level l's current iterate is in xs
  unsigned par = (A.phase == 2 && (max(A.pre - 1, 0) & 1)) ? (1u << (A.n - 1)) - 1u : 0u;
  auto cur = [&](int l) { return ((par >> l) & 1u) ? A.xs[l] : A.x[l]; };
  auto oth = [&](int l) { return ((par >> l) & 1u) ? A.x[l] : A.xs[l]; };
  // one tile pass over level l (every thread runs the same number of tiles)
  auto pass = [&](int l, auto&& op) {
    const LevelView<T>& L = A.L[l];
    const int groups = (L.nblocks + G::BPW - 1) / G::BPW;
    for (int g0 = 0; g0 < groups; g0 += nq) {
      const Place<BX, BY> P(g0 + quarter, L.nblocks, t, quarter, nq);
      op(L, P);
    }
    __syncthreads();
  };
  for (int l = 0; l + 1 < A.n && A.phase != 2; ++l) {
    for (int q = 1; q < A.pre; ++q) {
      T* xi = cur(l);
      T* xo = oth(l);
      pass(l, [&](const LevelView<T>& L, const Place<BX, BY>& P) {
        sweep_op<T, BX, BY, false, double>(S_, L, P, Arr4<T>{xi}, xo, A.b[l], A.omega[l], nullptr, nullptr);
      });
      par ^= 1u << l;
    }
    T* xl = cur(l);
    pass(l, [&](const LevelView<T>& L, const Place<BX, BY>& P) {
      residual_op<T, BX, BY>(S_, L, P, xl, A.b[l], A.u[l], A.omega_p[l]);
    });
    const int zero_start = (l + 2 < A.n && A.pre > 0) ? 1 : 0;
    pass(l, [&](const LevelView<T>& L, const Place<BX, BY>& P) {
      restrict_op<T, BX, BY>(S_, L, P, A.u[l], A.omega_p[l], A.b[l + 1], A.x[l + 1], A.dinv[l + 1], A.omega[l + 1],
                             zero_start);
    });
  }
  if (A.phase == 1) return;
  if (A.phase == 0) {
    coarse_op<T>(A.coarse_n, A.inv, A.b[A.n - 1], A.x[A.n - 1]);
    __syncthreads();
  }
  for (int l = A.n - 2; l >= 0; --l) {
    T* xl = cur(l);
    const T* xc = cur(l + 1);
    pass(l, [&](const LevelView<T>& L, const Place<BX, BY>& P) {
      prolong_op<T, BX, BY>(S_, L, P, xl, xc, A.omega_p[l]);
    });
    for (int q = 0; q < A.post; ++q) {
      T* xi = cur(l);
      T* xo = oth(l);
      pass(l, [&](const LevelView<T>& L, const Place<BX, BY>& P) {
        sweep_op<T, BX, BY, false, double>(S_, L, P, Arr4<T>{xi}, xo, A.b[l], A.omega[l], nullptr, nullptr);
      });
      par ^= 1u << l;
    }
  }
}

// y = A p with per-block partial p.y accumulated in FP64 (FP64 Krylov
// operator, or the FP32 finest level of the inner Krylov iteration)
// SUM: also the per-block sums of y (part[nb + b]).
template <typename T, int BX, int BY, bool SUM>
__global__ void __launch_bounds__(256, TILE_WAVES(T)) k_matvec_dot(LevelView<T> L, const T* __restrict__ p, T* __restrict__ y,
                                                   double* __restrict__ part, const int* __restrict__ done) {
  if (done && *done) return;
  TILE4_SHARED(T, BX, BY);
  const int wg = xcd_chunk_group(blockIdx.x, L.wg_chunk);
  if (wg * Geo<BX, BY>::BPW >= L.nblocks) return;  // grid padding
  const Place<BX, BY> P(wg, L.nblocks, threadIdx.x, 0);
  Vec4<T> vc, ax, d;
  tile4<T, BX, BY, false>(S_, L, P, Arr4<T>{p}, Pre0{}, vc, ax, d);
  double s = 0, sy = 0;
  if (P.valid) {
    st4(y + P.cell0, ax);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      s += double(vc.v[m]) * double(ax.v[m]);
      if constexpr (SUM) sy += double(ax.v[m]);
    }
  }
  if (part) {
    __shared__ double red[4];
    s = block4_sum<double, BX, BY>(s, red);
    if (P.valid && P.lt == 0) part[P.b] = s;
    if constexpr (SUM) {
      sy = block4_sum<double, BX, BY>(sy, red);
      if (P.valid && P.lt == 0) part[L.nblocks + P.b] = sy;
    }
  }
}

// FP64 residual replacement of the inner FP32 Krylov iteration (runs when
// the state requests one): r32 = scale (b - A x) with the FP64 operator, the
// V-cycle's zero-start iterate x0 = w D^-1 r32 and the workgroup's maximum of
// |r32| / V (a workgroup holds 1024 cells, the grouping of k_ir_update_xr);
// x0 = NULL: the zero-start iterate is left to the V-cycle's first sweep.
template <int BX, int BY>
__global__ void __launch_bounds__(256) k_ir_replace(LevelView<double> L, const double* __restrict__ x,
                                                   const double* __restrict__ rhs, float* __restrict__ r32,
                                                   float* __restrict__ x0, const float* __restrict__ dinv,
                                                   float omega, double* __restrict__ maxpart, CgState* st) {
  if (st->done || !st->replace) return;
  TILE4_SHARED(double, BX, BY);
  const int wg = xcd_chunk_group(blockIdx.x, L.wg_chunk);
  if (wg * Geo<BX, BY>::BPW >= L.nblocks) return;  // grid padding
  const Place<BX, BY> P(wg, L.nblocks, threadIdx.x, 0);
  Vec4<double> bv{}, vc, ax, d;
  Vec4<float> dv{};
  tile4<double, BX, BY, false>(S_, L, P, Arr4<double>{x}, [&]() {
    bv = ld4(rhs + P.cell0);
    if (x0) dv = ld4(dinv + P.cell0);
  }, vc, ax, d);
  double q = 0;
  if (P.valid) {
    const double sc = st->scale;
    Vec4<float> rv, xv;
    bool nan = false;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      rv.v[m] = float(sc * (bv.v[m] - ax.v[m]));
      xv.v[m] = omega * dv.v[m] * rv.v[m];
      q = fmax(q, double(fabsf(rv.v[m])));
      nan = nan || rv.v[m] != rv.v[m];
    }
    st4(r32 + P.cell0, rv);
    if (x0) st4(x0 + P.cell0, xv);
    q /= L.volume[P.b];
    if (nan) q = __longlong_as_double(0x7ff8000000000000LL);
  }
  __shared__ double red[4];
  q = block_max<double, 256>(q, red);
  if (threadIdx.x == 0) maxpart[wg] = q;
}

// Factorized round-off bound per cell and the per-block maximum of
// max(bound / V, |b| / V).
template <int BX, int BY>
__global__ void __launch_bounds__(256) k_floor_bound(LevelView<double> L, const double* __restrict__ x,
                                                    const double* __restrict__ rhs, double* __restrict__ part) {
  TILE4_SHARED(double, BX, BY);
  const int wg = xcd_chunk_group(blockIdx.x, L.wg_chunk);
  if (wg * Geo<BX, BY>::BPW >= L.nblocks) return;  // grid padding
  const Place<BX, BY> P(wg, L.nblocks, threadIdx.x, 0);
  Vec4<double> bv{}, vc, ax, d;
  tile4<double, BX, BY, true>(S_, L, P, Arr4<double>{x}, [&]() { bv = ld4(rhs + P.cell0); }, vc, ax, d);
  double qm = 0;
  if (P.valid) {
    const double V = L.volume[P.b];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      double q = fmax(ax.v[m] / V, fabs(bv.v[m]) / V);
      if (q != q) q = __longlong_as_double(0x7ff8000000000000LL);
      qm = (q != q || qm != qm) ? q : fmax(qm, q);
    }
  }
  __shared__ double red[4];
  qm = block4_max<double, BX, BY>(qm, red);
  if (P.valid && P.lt == 0) part[P.b] = qm;
}
