// Plane-marching fused smoothing passes of the 3D constant-coefficient
// multigrid levels (gfx950).
//
// A V-cycle level does, per pass of nu = 2 red-black sweeps, four colour
// stages (red, black, red, black).  As separate launches every stage
// streams the whole correction through HBM (half of each cache line is
// written, the line goes back whole), and the residual/restriction and the
// prolongation are passes of their own.  Here one workgroup owns a tile of
// 64 x 16 cells in (x, y) and a range of z planes, and marches up z with the
// four stages one plane apart: at step L
//     pre  pass:  stage 1 on plane L (from x = 0), stage s on plane L-s+1,
//                 residual on plane L-4 (restriction one step later);
//     post pass:  load plane L (x + prolongated coarse correction),
//                 stage s on plane L-s, write plane L-4.
// The tile is grown by a 4-cell halo in x and y, recomputed redundantly: a
// stage updates the tile grown by (4 - s) (pre: 5 - s), which is what the
// next stage reads, so every interior value equals the separate launches'
// value bit for bit (the same operations in the same order: stencil_core,
// the sums of mg_prolong_kernel and mg_resid_restrict_kernel).
//
// Each thread owns an x-pair (two cells) of the grown tile in every plane
// and keeps its values of the six planes in flight in registers (its rhs
// in a thread-private LDS ring): a cell's z neighbours and its pair mate
// are the thread's own.  Only in-plane neighbours of other threads go through LDS, one
// plane per slot of a 5-plane ring, each row stored colour-compacted (even
// x then odd x) so a wavefront's reads are consecutive words.  A stage's
// in-plane neighbours on plane q were written by the previous stage on
// plane q one step earlier, and no stage of a step writes a cell another
// stage of that step reads, so one barrier per step suffices.
//
// Memory per pass per component, in units of the level's array: pre reads
// f once (plus halo) and writes x and f/8; post reads x, f and x_coarse/8
// and writes x once — against about 2.6 units per colour stage, plus the
// residual/restriction and prolongation passes, of the separate kernels.
#ifndef VD_MG_FUSED_H_
#define VD_MG_FUSED_H_

namespace vd {

constexpr int kFTX = 64, kFTY = 16, kFH = 4;                    // tile, halo
constexpr int kFRX = kFTX + 2 * kFH, kFRY = kFTY + 2 * kFH;     // region 72 x 24
constexpr int kFHX = kFRX / 2;                                  // pairs per row
constexpr int kFPS = kFRX * kFRY;                               // region plane
constexpr int kFNS = 5;                                         // LDS ring planes
constexpr int kFPairs = kFHX * kFRY;                            // pairs per plane
// post pass: coarse window of the region's parents and their neighbours
// (38 x 14 per coarse plane), staged in an LDS ring of 3 coarse planes
constexpr int kCSX = kFRX / 2 + 2, kCSY = kFRY / 2 + 2, kCS = kCSX * kCSY;
// x-pairs per thread (pair p = tid + u * kFThreads).  One pair per thread
// and 14 waves measured faster than two pairs and 7 waves (which avoid
// register spills in FP64 but hide less latency).
constexpr int kFPPT = 1;
constexpr int kFThreads = 896;
static_assert(kFPairs <= kFPPT * kFThreads, "x-pairs per workgroup");
static_assert(kCS <= kFThreads, "one staged coarse value per thread");
constexpr int kFFW = kFPPT * kFThreads;                         // rhs ring row (one word per pair)
static_assert(kFH % 2 == 0 && kFTX % 2 == 0 && kFTY % 2 == 0, "pairs and children aligned to even x and y");

template <class T>
struct FusedArgs
{
    MGView<T> v;                 // this level
    const LevelBC* bcs;          // this level's boundary model
    int pm;                      // periodic mask
    const T* f;                  // rhs, nc components
    const T* xin;                // post: pre-smoothed correction (null: zero)
    MGView<T> vc;                // post: coarse level
    const T* xcg;                // post: coarse correction with one ghost layer
                                 // ((c0+2) x (c1+2) x (c2+2) per component);
                                 // null: no prolongation (a refined level)
    T* xout;                     // correction after the pass
    T* fc;                       // pre: coarse rhs (restricted residual)
    T* rout;                     // pre, refined level: the residual itself (no
                                 // restriction)
    long long Nc;                // pre: coarse cells per component
    int cn0, cn1;                // pre: coarse extents in x, y
    int zc, nchunk;              // z planes per workgroup, chunks per component
    unsigned active;
    // the active components in order: grid z enumerates (component slot,
    // chunk), so inactive components get no workgroups (an early-exiting
    // workgroup still holds a CU's LDS while it is dispatched, and those
    // delayed the active ones by up to the pass time again)
    int comp[kMaxComp];
    // post on the finest level: add the result to the solution instead of
    // writing xout (component c at sol + compmap[c] * sstride; null: xout)
    double* sol;
    long long sstride;
    const int* compmap;
};

// LDS ring slot of plane p (p may be negative)
__device__ __forceinline__ int fz_slot(int p)
{
    const int r = p % kFNS;
    return r < 0 ? r + kFNS : r;
}
// Within a plane, region cell (ra, rb) sits colour-compacted at
// rb * kFRX + (ra & 1) * kFHX + (ra >> 1).

// Per-workgroup geometry shared by both passes.
struct FusedTile
{
    int n0, n1, n2;
    bool px, py, pz;
    int x0, y0, ix1, iy1, z0, z1;   // interior box (x, y clipped) and z range
    int rx0, ry0;                   // region origin (global, unwrapped)
    __device__ bool alive_x(int gx) const { return px || (gx >= 0 && gx < n0); }
    __device__ bool alive_y(int gy) const { return py || (gy >= 0 && gy < n1); }
    __device__ bool alive_z(int gz) const { return pz || (gz >= 0 && gz < n2); }
    // cells a pass loads: in the domain (or a periodic image) and within the
    // clipped interior grown by the halo, the most any stage reads; this also
    // keeps a periodic image within one period of the domain (n >= kFH)
    __device__ bool load_x(int gx) const { return alive_x(gx) && gx >= x0 - kFH && gx < ix1 + kFH; }
    __device__ bool load_y(int gy) const { return alive_y(gy) && gy >= y0 - kFH && gy < iy1 + kFH; }
    __device__ static int wrap(int g, int n) { return g < 0 ? g + n : (g >= n ? g - n : g); }
    __device__ long long cell(int gx, int gy, int gz) const
    {
        return (long long)wrap(gx, n0) + (long long)n0 * ((long long)wrap(gy, n1) + (long long)n1 * wrap(gz, n2));
    }
};

template <class T>
__device__ __forceinline__ FusedTile fused_tile(const FusedArgs<T>& A, const dim3& b)
{
    FusedTile t;
    t.n0 = A.v.n[0];
    t.n1 = A.v.n[1];
    t.n2 = A.v.n[2];
    t.px = A.pm & 1;
    t.py = (A.pm >> 1) & 1;
    t.pz = (A.pm >> 2) & 1;
    t.x0 = (int)b.x * kFTX;
    t.y0 = (int)b.y * kFTY;
    t.ix1 = min(t.x0 + kFTX, t.n0);
    t.iy1 = min(t.y0 + kFTY, t.n1);
    const int chunk = (int)b.z % A.nchunk;
    t.z0 = chunk * A.zc;
    t.z1 = min(t.z0 + A.zc, t.n2);
    t.rx0 = t.x0 - kFH;
    t.ry0 = t.y0 - kFH;
    return t;
}

// Thread-owned data: the x-pair (cells ra0, ra0+1 of region row rb) of the
// six planes L-k, k = 0..5 (k = 0 newest), and their rhs (five planes, in
// LDS), held by colour: r = the pair's colour-0 cell, b = its colour-1 cell.  A stage
// updates one colour, its z neighbours have the other colour, so every
// register access is fixed at compile time; only the x position of a
// colour, (gy + q + colour) & 1, varies from plane to plane.
template <class T>
struct FusedCol
{
    T r[6], b[6];
    T* fl;                               // the pair's rhs in LDS: fl[(slot * 2 + colour) * kFFW]
    int ra0, rb, gx0, gy;
    int cb;                              // colour-compacted index of the pair's even cell
    int de, dn;                          // distance of the even / odd cell outside the
                                         // tile in x, y (max norm; large if not alive)
    bool lo0, hi0, lo1, hi1, loy, hiy;   // non-periodic domain sides of the two cells
    T dg0, dg1;                          // diagonals of the two cells off the z sides
    T ig0, ig1;                          // their reciprocals

    __device__ __forceinline__ void shift()
    {
#pragma unroll
        for (int k = 5; k > 0; --k)
        {
            r[k] = r[k - 1];
            b[k] = b[k - 1];
        }
    }
    // rhs of the pair's colour-col cell on plane q (q in flight)
    __device__ __forceinline__ T f(int q, int col) const { return fl[(fz_slot(q) * 2 + col) * kFFW]; }
    // new plane L from the x-even / x-odd values
    __device__ __forceinline__ void put(int L, T xe, T xo, T fe, T fo)
    {
        const bool sw = (gy + L) & 1;   // colour 0 sits at the odd x
        r[0] = sw ? xo : xe;
        b[0] = sw ? xe : xo;
        T* fs = fl + fz_slot(L) * 2 * kFFW;
        fs[0] = sw ? fo : fe;
        fs[kFFW] = sw ? fe : fo;
    }
};

// The cell a step's stages update.  Every stage of a step updates, in each
// pair, the cell at the same x position: in the pre pass stage s works on
// plane L-s with colour s&1 (the from-zero stage is s = 0), in the post pass
// on plane L-s with colour (s-1)&1, so (gy + plane + colour) & 1 is the same
// for all of them.  Its LDS slots are therefore fixed for the step: self at
// cb + off * kFHX, its y neighbours one row (kFRX) away, its x neighbour in
// the adjacent pair (kFHX - 1) away -- the left pair's odd cell for an even
// cell, the right pair's even cell for an odd one.
struct FStep
{
    int off;      // x position in the pair (0 even, 1 odd)
    int self;     // colour-compacted index within a plane
    int across;   // index of the x neighbour in the adjacent pair
    int d;        // distance outside the tile (x, y), large if not alive
};
template <class T>
__device__ __forceinline__ FStep fused_step(const FusedCol<T>& C, int off)
{
    FStep S;
    S.off = off;
    S.self = C.cb + (off ? kFHX : 0);
    S.across = S.self + (off ? -(kFHX - 1) : (kFHX - 1));
    S.d = off ? C.dn : C.de;
    return S;
}

// Operator row of the thread's cell S.off on plane q, of colour COL, held
// at register slot K: x neighbours = pair mate and the neighbouring pair's
// cell (LDS), y neighbours from LDS, z neighbours (the other colour) from
// slots K-1 (plane q+1) and K+1 (plane q-1).  idg is the reciprocal of the
// diagonal: off the z sides the precomputed one (of the value stencil_core
// gives there).  Extents are
// >= 8 here, so a neighbour replaced at a domain side in stencil_b never
// reaches the result and is not replaced.  ws: some lane of the wave owns a
// cell on an x or y domain side (uniform); without it the side flags are
// known false and stencil_core selects no ghost (same values, same
// operations in the same order).
template <class T, int K, int COL>
__device__ __forceinline__ void fused_row(const MGView<T>& v, const SideBC (*bc)[2], const FusedTile& t,
                                          const FusedCol<T>& C, const T* __restrict__ plane, int q, T bcoef,
                                          const FStep& S, bool ws, T& xc, T& Lx, T& idg)
{
    const int off = S.off;
    xc = COL ? C.b[K] : C.r[K];
    const T mate = COL ? C.r[K] : C.b[K];
    const T across = plane[S.across];
    T xmr[3], xpr[3];
    xmr[0] = off ? mate : across;
    xpr[0] = off ? across : mate;
    xmr[1] = plane[S.self - kFRX];
    xpr[1] = plane[S.self + kFRX];
    xmr[2] = COL ? C.r[K + 1] : C.b[K + 1];
    xpr[2] = COL ? C.r[K - 1] : C.b[K - 1];
    const bool zlo = !t.pz && q == 0, zhi = !t.pz && q == t.n2 - 1;   // uniform
    const T bm[3] = {bcoef, bcoef, bcoef};
    if (zlo || zhi)
    {
        // the diagonal is evaluated here: always from the lane's own flags,
        // so its rounding is that of the per-colour kernels
        bool lob[3], hib[3];
        lob[0] = off ? C.lo1 : C.lo0;
        hib[0] = off ? C.hi1 : C.hi0;
        lob[1] = C.loy;
        hib[1] = C.hiy;
        lob[2] = zlo;
        hib[2] = zhi;
        T dg;
        stencil_core<3, T, true>(v, bc, xc, xmr, xpr, lob, hib, bm, bm, Lx, dg);
        idg = T(1) / dg;
        return;
    }
    T dd;
    if (ws)
    {
        bool lob[3], hib[3];
        lob[0] = off ? C.lo1 : C.lo0;
        hib[0] = off ? C.hi1 : C.hi0;
        lob[1] = C.loy;
        hib[1] = C.hiy;
        lob[2] = hib[2] = false;
        stencil_core<3, T, true>(v, bc, xc, xmr, xpr, lob, hib, bm, bm, Lx, dd);
    } else
    {
        const bool lob[3] = {false, false, false}, hib[3] = {false, false, false};
        stencil_core<3, T, true>(v, bc, xc, xmr, xpr, lob, hib, bm, bm, Lx, dd);
    }
    idg = off ? C.ig1 : C.ig0;
}

// Colour-COL stage on plane q (register slot K): the thread's cell of that
// colour (S.off) is updated when it lies in the interior grown by g.
template <class T, int K, int COL>
__device__ __forceinline__ void fused_stage(const MGView<T>& v, const SideBC (*bc)[2], const FusedTile& t,
                                            FusedCol<T>& C, T* ring, int q, int g, T bcoef, const FStep& S, bool ws)
{
    if (q < t.z0 - g || q >= t.z1 + g || !t.alive_z(q)) return;
    if (S.d > g) return;
    T* plane = ring + fz_slot(q) * kFPS;
    T xc, Lx, idg;
    fused_row<T, K, COL>(v, bc, t, C, plane, q, bcoef, S, ws, xc, Lx, idg);
    const T fv = C.f(q, COL);
    const T xn = gs_update(xc, fv - Lx, idg);
    if constexpr (COL) C.b[K] = xn;
    else C.r[K] = xn;
    plane[S.self] = xn;
}

// Pair of lane p.  Pairs with a cell on an x or y domain side (non-periodic)
// take the last lanes, so at most the workgroup's last two waves hold them
// and the others run the operator rows without ghost selection (fused_row,
// ws).  Side columns: even cells at x = 0 and odd cells at x = n0 - 1 (gx0 is
// even, n0 even), side rows: y = 0 and y = n1 - 1; non-side pairs keep
// row-major order, so a wave still reads consecutive LDS words.
__device__ __forceinline__ void fused_pair_of(const FusedTile& t, int p, int& hc, int& rb)
{
    int sc0 = 0, sc1 = 0, nsc = 0, sr0 = 0, sr1 = 0, nsr = 0;   // ascending
    if (!t.px)
    {
        const int a = (0 - t.rx0) >> 1, b = (t.n0 - 2 - t.rx0) >> 1;
        if (a >= 0 && a < kFHX) sc0 = a, nsc = 1;
        if (b >= 0 && b < kFHX)
        {
            if (nsc) sc1 = b;
            else sc0 = b;
            ++nsc;
        }
    }
    if (!t.py)
    {
        const int a = 0 - t.ry0, b = t.n1 - 1 - t.ry0;
        if (a >= 0 && a < kFRY) sr0 = a, nsr = 1;
        if (b >= 0 && b < kFRY)
        {
            if (nsr) sr1 = b;
            else sr0 = b;
            ++nsr;
        }
    }
    // i-th value not excluded by the (ascending) exclusions e0, e1
    auto skip = [](int i, int e0, int e1, int ne) __attribute__((always_inline)) {
        if (ne > 0 && i >= e0) ++i;
        if (ne > 1 && i >= e1) ++i;
        return i;
    };
    const int ncn = kFHX - nsc, nrn = kFRY - nsr, nn = ncn * nrn;
    if (p < nn)
    {
        rb = skip(p / ncn, sr0, sr1, nsr);
        hc = skip(p % ncn, sc0, sc1, nsc);
        return;
    }
    int q = p - nn;
    if (q < nsr * kFHX)
    {
        rb = q < kFHX ? sr0 : sr1;
        hc = q % kFHX;
        return;
    }
    q -= nsr * kFHX;
    hc = q < nrn ? sc0 : sc1;
    rb = skip(q % nrn, sr0, sr1, nsr);
}

// Thread setup shared by both passes: pair position, side flags, distances
// outside the tile and the diagonals of its two cells off the z sides.
// Returns whether the pair has a cell on an x or y domain side.
template <class T>
__device__ __forceinline__ bool fused_col_init(const MGView<T>& v, const SideBC (*bc)[2], const FusedTile& t,
                                               FusedCol<T>& C, int p, T bcoef, T* fl)
{
    C.fl = fl;
    int hc, rb;
    fused_pair_of(t, p, hc, rb);
    C.ra0 = 2 * hc;
    C.rb = rb;
    C.cb = rb * kFRX + hc;
    C.gx0 = t.rx0 + C.ra0;
    C.gy = t.ry0 + C.rb;
    C.lo0 = !t.px && C.gx0 == 0;
    C.hi0 = !t.px && C.gx0 == t.n0 - 1;
    C.lo1 = !t.px && C.gx0 + 1 == 0;
    C.hi1 = !t.px && C.gx0 + 1 == t.n0 - 1;
    C.loy = !t.py && C.gy == 0;
    C.hiy = !t.py && C.gy == t.n1 - 1;
    auto dist = [](int g, int lo, int hi, bool alive) __attribute__((always_inline)) {
        return alive ? max(max(lo - g, g - (hi - 1)), 0) : (1 << 20);
    };
    const int dy = dist(C.gy, t.y0, t.iy1, t.alive_y(C.gy));
    C.de = max(dist(C.gx0, t.x0, t.ix1, t.alive_x(C.gx0)), dy);
    C.dn = max(dist(C.gx0 + 1, t.x0, t.ix1, t.alive_x(C.gx0 + 1)), dy);
    const T z[3] = {T(0), T(0), T(0)}, bm[3] = {bcoef, bcoef, bcoef};
    T Lx;
    {
        const bool lob[3] = {C.lo0, C.loy, false}, hib[3] = {C.hi0, C.hiy, false};
        stencil_core<3, T, true>(v, bc, T(0), z, z, lob, hib, bm, bm, Lx, C.dg0);
    }
    {
        const bool lob[3] = {C.lo1, C.loy, false}, hib[3] = {C.hi1, C.hiy, false};
        stencil_core<3, T, true>(v, bc, T(0), z, z, lob, hib, bm, bm, Lx, C.dg1);
    }
    C.ig0 = T(1) / C.dg0;
    C.ig1 = T(1) / C.dg1;
#pragma unroll
    for (int k = 0; k < 6; ++k) C.r[k] = C.b[k] = T(0);
    return C.lo0 || C.hi0 || C.lo1 || C.hi1 || C.loy || C.hiy;
}

// Pre-smoothing pass from x = 0 (nu1 = 2) fused with residual + restriction:
// writes the correction (interior) and the coarse rhs.  The residuals of a
// coarse cell's odd fine row go through LDS (double-buffered by step
// parity); the thread of its even row sums its children one step later, in
// the order of mg_resid_restrict_kernel.
template <class T>
__global__ void __launch_bounds__(kFThreads) mg_fused_pre_kernel(FusedArgs<T> A)
{
    __shared__ T ring[kFNS * kFPS];
    __shared__ T fring[kFNS * 2 * kFFW];
    __shared__ T rodd[2][kFTX * (kFTY / 2)];
    const dim3 b = xcd_block();
    const int c = __builtin_amdgcn_readfirstlane(A.comp[(int)b.z / A.nchunk]);
    const FusedTile t = fused_tile(A, b);
    const MGView<T>& v = A.v;
    // the side models of this component, read once (uniform)
    SideBC bc[3][2];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) bc[d][s2] = A.bcs->s[c][d][s2];
    const T bcoef = T(v.bcomp ? v.bcomp[c] : v.bconst);
    const long long cb = (long long)c * v.Nl;
    const T* __restrict__ f = A.f + cb;
    const int tid = threadIdx.x;
    FusedCol<T> C[kFPPT];
    bool pair[kFPPT], inter[kFPPT], owner[kFPPT];
    int ex[kFPPT], ey[kFPPT];
    T rsum[kFPPT], r00[kFPPT], r01[kFPPT], fn0[kFPPT], fn1[kFPPT];
    bool side = false;
#pragma unroll
    for (int u = 0; u < kFPPT; ++u)
    {
        const int p = tid + u * kFThreads;
        pair[u] = p < kFPairs;
        side |= fused_col_init<T>(v, bc, t, C[u], pair[u] ? p : 0, bcoef, fring + p) && pair[u];
        // interior pair: residuals and the write-out; an even interior row
        // owns the coarse cells over its children
        inter[u] = pair[u] && C[u].gy >= t.y0 && C[u].gy < t.iy1 && C[u].gx0 >= t.x0 && C[u].gx0 < t.ix1;
        ex[u] = C[u].gx0 - t.x0;
        ey[u] = C[u].gy - t.y0;
        owner[u] = inter[u] && (ey[u] & 1) == 0 && !A.rout;
        rsum[u] = r00[u] = r01[u] = fn0[u] = fn1[u] = T(0);
    }
    const bool ws = __builtin_amdgcn_ballot_w64(side) != 0;   // uniform per wave
    // restriction of plane q: own children (kept) + the odd row's (LDS), in
    // the order (jj, ii) of mg_resid_restrict_kernel
    auto restrict_plane = [&](int u, int q, const T* ro) __attribute__((always_inline)) {
        if (!owner[u] || q < t.z0 || q >= t.z1) return;
        ro += (ey[u] >> 1) * kFTX + ex[u];
        if (((q - t.z0) & 1) == 0) rsum[u] = T(0);
        rsum[u] += r00[u];
        rsum[u] += r01[u];
        rsum[u] += ro[0];
        rsum[u] += ro[1];
        if ((q - t.z0) & 1)
        {
            const int I = (t.x0 + ex[u]) >> 1, J = (t.y0 + ey[u]) >> 1, Kc = q >> 1;
            A.fc[(long long)c * A.Nc + (long long)I + (long long)A.cn0 * ((long long)J + (long long)A.cn1 * Kc)] =
                rsum[u] / T(8);
        }
    };
    auto fetch_f = [&](int u, int p) __attribute__((always_inline)) {
        fn0[u] = fn1[u] = T(0);
        if (!pair[u] || !t.alive_z(p) || !t.load_y(C[u].gy)) return;
        if (t.load_x(C[u].gx0)) fn0[u] = f[t.cell(C[u].gx0, C[u].gy, p)];
        if (t.load_x(C[u].gx0 + 1)) fn1[u] = f[t.cell(C[u].gx0 + 1, C[u].gy, p)];
    };
    const int Lb = t.z0 - kFH, Le = t.z1 + kFH;   // steps L in [Lb, Le)
#pragma unroll
    for (int u = 0; u < kFPPT; ++u) fetch_f(u, Lb);
    FStep S[kFPPT];
    for (int L = Lb; L < Le; ++L)
    {
#pragma unroll
        for (int u = 0; u < kFPPT; ++u)
        {
            restrict_plane(u, L - 5, rodd[(L - 1) & 1]);
            S[u] = fused_step(C[u], (C[u].gy + L) & 1);
            C[u].shift();
            C[u].put(L, T(0), T(0), fn0[u], fn1[u]);
            if (L + 1 < Le) fetch_f(u, L + 1);
        }
#pragma unroll
        for (int u = 0; u < kFPPT; ++u)
        {
            if (!pair[u]) continue;
            FusedCol<T>& Q = C[u];
            // stage 1 on plane L from x = 0: colour 0 gets f/diag, colour 1 zero
            const int off = S[u].off;   // x position of colour 0
            if (t.alive_z(L) && S[u].d <= kFH)
            {
                T idg = off ? Q.ig1 : Q.ig0;
                if (!t.pz && (L == 0 || L == t.n2 - 1))
                {
                    bool lob[3], hib[3];
                    lob[0] = off ? Q.lo1 : Q.lo0;
                    hib[0] = off ? Q.hi1 : Q.hi0;
                    lob[1] = Q.loy;
                    hib[1] = Q.hiy;
                    lob[2] = L == 0;
                    hib[2] = L == t.n2 - 1;
                    const T zr[3] = {T(0), T(0), T(0)}, bm[3] = {bcoef, bcoef, bcoef};
                    T Lx, dg;
                    stencil_core<3, T, true>(v, bc, T(0), zr, zr, lob, hib, bm, bm, Lx, dg);
                    idg = T(1) / dg;
                }
                Q.r[0] = gs_update(T(0), Q.f(L, 0) - T(0), idg);
            }
            T* pl = ring + fz_slot(L) * kFPS;
            pl[S[u].self] = Q.r[0];
            pl[Q.cb + (off ? 0 : kFHX)] = T(0);
        }
#pragma unroll
        for (int u = 0; u < kFPPT; ++u)
            if (pair[u]) fused_stage<T, 1, 1>(v, bc, t, C[u], ring, L - 1, 3, bcoef, S[u], ws);
#pragma unroll
        for (int u = 0; u < kFPPT; ++u)
            if (pair[u]) fused_stage<T, 2, 0>(v, bc, t, C[u], ring, L - 2, 2, bcoef, S[u], ws);
#pragma unroll
        for (int u = 0; u < kFPPT; ++u)
            if (pair[u]) fused_stage<T, 3, 1>(v, bc, t, C[u], ring, L - 3, 1, bcoef, S[u], ws);
        // plane L-4 is final: write it, residuals of the tile's cells
        const int q = L - 4;
#pragma unroll
        for (int u = 0; u < kFPPT; ++u)
        {
            if (!inter[u] || q < t.z0 || q >= t.z1) continue;
            const FusedCol<T>& Q = C[u];
            const bool sw = S[u].off;   // colour 0 at the odd x: (gy + q) & 1
            const long long o = cb + (long long)Q.gx0 + (long long)t.n0 * ((long long)Q.gy + (long long)t.n1 * q);
            A.xout[o] = sw ? Q.b[4] : Q.r[4];
            A.xout[o + 1] = sw ? Q.r[4] : Q.b[4];
            const T* plane = ring + fz_slot(q) * kFPS;
            T xc, Lx, dg;
            fused_row<T, 4, 0>(v, bc, t, Q, plane, q, bcoef, S[u], ws, xc, Lx, dg);
            const T rr = Q.f(q, 0) - Lx;
            fused_row<T, 4, 1>(v, bc, t, Q, plane, q, bcoef, fused_step(Q, 1 - S[u].off), ws, xc, Lx, dg);
            const T rbk = Q.f(q, 1) - Lx;
            const T re = sw ? rbk : rr, rodv = sw ? rr : rbk;   // x-even, x-odd cell
            if (A.rout)
            {
                A.rout[o] = re;
                A.rout[o + 1] = rodv;
            } else if (owner[u])
            {
                r00[u] = re;
                r01[u] = rodv;
            } else
            {
                T* ro = rodd[L & 1] + (ey[u] >> 1) * kFTX + ex[u];
                ro[0] = re;
                ro[1] = rodv;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < kFPPT; ++u) restrict_plane(u, Le - 5, rodd[(Le - 1) & 1]);
}

// Post-smoothing pass (nu2 = 2) fused with the prolongation of the coarse
// correction: x_out = smooth(x_in + P x_coarse) on the interior.
template <class T>
__global__ void __launch_bounds__(kFThreads) mg_fused_post_kernel(FusedArgs<T> A)
{
    __shared__ T ring[kFNS * kFPS];
    __shared__ T fring[kFNS * 2 * kFFW];
    __shared__ T cst[3 * kCS];
    const dim3 b = xcd_block();
    const int c = __builtin_amdgcn_readfirstlane(A.comp[(int)b.z / A.nchunk]);
    const FusedTile t = fused_tile(A, b);
    const MGView<T>& v = A.v;
    SideBC bc[3][2];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) bc[d][s2] = A.bcs->s[c][d][s2];
    const T bcoef = T(v.bcomp ? v.bcomp[c] : v.bconst);
    const long long cb = (long long)c * v.Nl;
    const T* __restrict__ f = A.f + cb;
    // no xin: the pass starts from x = 0 (smoothing iterations of the
    // solve, which have no coarse correction either)
    const T* __restrict__ xin = A.xin ? A.xin + cb : nullptr;
    const int g0 = A.vc.n[0] + 2, g1 = A.vc.n[1] + 2;
    const T* __restrict__ xcb = A.xcg ? A.xcg + (long long)c * g0 * g1 * (A.vc.n[2] + 2) : nullptr;
    const int tid = threadIdx.x;
    // Staged coarse window: entry (sc, sr) of a plane holds the ghosted
    // array's value at unwrapped coarse index (bxc + sc, byc + sr, U): the
    // cell or its periodic image, or the ghost layer beyond a non-periodic
    // side (indices clamped to that layer where no live pair reads them) --
    // the values a pair would read from the ghosted array directly.
    const int bxc = (t.rx0 >> 1) - 1, byc = (t.ry0 >> 1) - 1;
    const bool stager = A.xcg != nullptr && tid < kCS;
    auto gidx = [](int U, int n, bool per) __attribute__((always_inline)) {
        if (per)
        {
            const int m = U % n;
            return (m < 0 ? m + n : m) + 1;
        }
        return min(max(U + 1, 0), n + 1);
    };
    const long long sxy = stager ? (long long)gidx(bxc + tid % kCSX, A.vc.n[0], t.px) +
                                       (long long)g0 * gidx(byc + tid / kCSX, A.vc.n[1], t.py)
                                 : 0;
    auto stage_load = [&](int U) __attribute__((always_inline)) {
        return xcb[sxy + (long long)g0 * g1 * gidx(U, A.vc.n[2], t.pz)];
    };
    auto cslot = [](int U) __attribute__((always_inline)) {
        const int m = U % 3;
        return m < 0 ? m + 3 : m;
    };
    FusedCol<T> C[kFPPT];
    bool pair[kFPPT], inter[kFPPT];
    bool side = false;
#pragma unroll
    for (int u = 0; u < kFPPT; ++u)
    {
        const int p = tid + u * kFThreads;
        pair[u] = p < kFPairs;
        side |= fused_col_init<T>(v, bc, t, C[u], pair[u] ? p : 0, bcoef, fring + p) && pair[u];
        inter[u] = pair[u] && C[u].gy >= t.y0 && C[u].gy < t.iy1 && C[u].gx0 >= t.x0 && C[u].gx0 < t.ix1;
    }
    const bool ws = __builtin_amdgcn_ballot_w64(side) != 0;   // uniform per wave
    // prefetched raw values of the next plane: a pair's x and f
    T xn0[kFPPT], xn1[kFPPT], fn0[kFPPT], fn1[kFPPT];
    bool live[kFPPT];
    auto fetch = [&](int u, int p) __attribute__((always_inline)) {
        const FusedCol<T>& Q = C[u];
        live[u] = pair[u] && t.alive_z(p) && t.load_y(Q.gy) && t.load_x(Q.gx0);
        if (!live[u]) return;
        const int i = FusedTile::wrap(Q.gx0, t.n0);
        const int j = FusedTile::wrap(Q.gy, t.n1), k = FusedTile::wrap(p, t.n2);
        const long long o = (long long)i + (long long)t.n0 * ((long long)j + (long long)t.n1 * k);
        xn0[u] = xin ? xin[o] : T(0);
        xn1[u] = xin ? xin[o + 1] : T(0);
        fn0[u] = f[o];
        fn1[u] = f[o + 1];
    };
    // the two x-children (x-even, x-odd) of the pair's parent on plane p:
    // the weighted sums of mg_prolong_kernel, in its order (cc, bb, aa),
    // neighbour I + aa*sx with sx = -1 (ii = 0) or +1.  The parent is (I,
    // J, K) in unwrapped coarse indices (n even: a periodic image keeps the
    // parity), its neighbours I-1..I+1, J+sy, K+sz come from the staged
    // window, read in the step that uses them (the staging schedule keeps
    // the planes of plane L in place through step L).
    auto children = [&](int u, int p, T& ce, T& co) __attribute__((always_inline)) {
        const FusedCol<T>& Q = C[u];
        const int I = Q.gx0 >> 1, J = Q.gy >> 1, Kp = p >> 1;
        const int sy = (Q.gy & 1) ? 1 : -1, sz = (p & 1) ? 1 : -1;
        T nb[2][2][3];
#pragma unroll
        for (int cc = 0; cc < 2; ++cc)
        {
            const T* pl = cst + cslot(Kp + cc * sz) * kCS + (I - 1 - bxc);
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int a = 0; a < 3; ++a) nb[cc][bb][a] = pl[(J + bb * sy - byc) * kCSX + a];
        }
#pragma unroll
        for (int ii = 0; ii < 2; ++ii)
        {
            const int sx = ii ? 1 : -1;
            T acc = T(0);
#pragma unroll
            for (int cc = 0; cc < 2; ++cc)
#pragma unroll
                for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                    for (int aa = 0; aa < 2; ++aa)
                    {
                        T w = aa ? T(0.25) : T(0.75);
                        w *= bb ? T(0.25) : T(0.75);
                        w *= cc ? T(0.25) : T(0.75);
                        acc += w * nb[cc][bb][1 + aa * sx];
                    }
            if (ii) co = acc;
            else ce = acc;
        }
    };
    const int Lb = t.z0 - kFH, Le = t.z1 + kFH;   // Lb even
    // coarse planes of fine planes Lb and Lb+1: K-1, K, K+1
    T stg = T(0);
    if (A.xcg)
    {
        if (stager)
            for (int U = (Lb >> 1) - 1; U <= (Lb >> 1) + 1; ++U) cst[cslot(U) * kCS + tid] = stage_load(U);
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < kFPPT; ++u) fetch(u, Lb);
    FStep S[kFPPT];
    for (int L = Lb; L < Le; ++L)
    {
        // an odd fine plane p first reads coarse plane (p >> 1) + 1: the
        // one of plane L+2 is written now (loaded a step ago), the one of
        // L+3 loaded; the slot written (coarse plane (L >> 1) + 2 for odd
        // L) is not one that plane L, read in this step, uses
        if (stager)
        {
            if ((L + 2) & 1) cst[cslot(((L + 2) >> 1) + 1) * kCS + tid] = stg;
            if ((L + 3) & 1) stg = stage_load(((L + 3) >> 1) + 1);
        }
#pragma unroll
        for (int u = 0; u < kFPPT; ++u)
        {
            FusedCol<T>& Q = C[u];
            S[u] = fused_step(Q, (Q.gy + L + 1) & 1);
            Q.shift();
            // with the prolongation x + P x_coarse, else x as it is
            T xe = T(0), xo = T(0);
            if (live[u])
            {
                xe = xn0[u];
                xo = xn1[u];
                if (A.xcg)
                {
                    T ce, co;
                    children(u, L, ce, co);
                    xe = xe + ce;
                    xo = xo + co;
                }
            }
            Q.put(L, xe, xo, live[u] ? fn0[u] : T(0), live[u] ? fn1[u] : T(0));
            if (pair[u])
            {
                T* pl = ring + fz_slot(L) * kFPS;
                pl[Q.cb] = xe;
                pl[Q.cb + kFHX] = xo;
            }
            if (L + 1 < Le) fetch(u, L + 1);
        }
#pragma unroll
        for (int u = 0; u < kFPPT; ++u)
            if (pair[u]) fused_stage<T, 1, 0>(v, bc, t, C[u], ring, L - 1, 3, bcoef, S[u], ws);
#pragma unroll
        for (int u = 0; u < kFPPT; ++u)
            if (pair[u]) fused_stage<T, 2, 1>(v, bc, t, C[u], ring, L - 2, 2, bcoef, S[u], ws);
#pragma unroll
        for (int u = 0; u < kFPPT; ++u)
            if (pair[u]) fused_stage<T, 3, 0>(v, bc, t, C[u], ring, L - 3, 1, bcoef, S[u], ws);
#pragma unroll
        for (int u = 0; u < kFPPT; ++u)
            if (pair[u]) fused_stage<T, 4, 1>(v, bc, t, C[u], ring, L - 4, 0, bcoef, S[u], ws);
        const int q = L - 4;
#pragma unroll
        for (int u = 0; u < kFPPT; ++u)
        {
            if (!inter[u] || q < t.z0 || q >= t.z1) continue;
            const FusedCol<T>& Q = C[u];
            const bool sw = (Q.gy + q) & 1;   // colour 0 at the odd x
            const T ve = sw ? Q.b[4] : Q.r[4], vo = sw ? Q.r[4] : Q.b[4];
            const long long oc = (long long)Q.gx0 + (long long)t.n0 * ((long long)Q.gy + (long long)t.n1 * q);
            if (A.sol)
            {
                // mg_add_kernel's update, x = x + e
                double* sp = A.sol + (long long)A.compmap[c] * A.sstride + oc;
                sp[0] = sp[0] + double(ve);
                sp[1] = sp[1] + double(vo);
            } else
            {
                A.xout[cb + oc] = ve;
                A.xout[cb + oc + 1] = vo;
            }
        }
        __syncthreads();
    }
}

// Coarse correction with one ghost layer for the prolongation: every entry
// is cval_x of its (possibly out-of-range) index, the value mg_prolong_kernel
// and prolong_value read there.
template <class T>
__global__ void __launch_bounds__(256) mg_coarse_ghost_kernel(MGView<T> vc, const LevelBC* __restrict__ bcs, int pm,
                                                              const T* __restrict__ xc, T* __restrict__ xg,
                                                              unsigned active)
{
    const int g0 = vc.n[0] + 2, g1 = vc.n[1] + 2, g2 = vc.n[2] + 2;
    const int c = (int)blockIdx.z / g2;
    const int K = (int)blockIdx.z - c * g2;
    if (!((active >> c) & 1u)) return;
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= g0 || J >= g1) return;
    const long long o = (long long)c * g0 * g1 * g2 + (long long)I + (long long)g0 * ((long long)J + (long long)g1 * K);
    const T* xcb = xc + (long long)c * vc.Nl;
    if (I >= 1 && I <= vc.n[0] && J >= 1 && J <= vc.n[1] && K >= 1 && K <= vc.n[2])
        xg[o] = xcb[(long long)(I - 1) + (long long)vc.n[0] * ((long long)(J - 1) + (long long)vc.n[1] * (K - 1))];
    else
        xg[o] = cval_x<3, T>(vc, bcs->s[c], pm, xcb, I - 1, J - 1, K - 1);
}

// x_f += P x_c from the ghosted coarse correction (3D): one thread per
// coarse cell, the 3x3x3 neighbourhood by direct loads, each child the sum
// of mg_prolong_kernel in its order.
template <class T>
__global__ void __launch_bounds__(256) mg_prolong_ghost_kernel(int nf0, int nf1, int nf2, long long Nf,
                                                               T* __restrict__ xf, MGView<T> vc,
                                                               const T* __restrict__ xg, unsigned active)
{
    const dim3 bI_ = xcd_block();
    const int c = bI_.z / vc.n[2];
    const int K = bI_.z - c * vc.n[2];
    if (!((active >> c) & 1u)) return;
    const int I = bI_.x * blockDim.x + threadIdx.x;
    const int J = bI_.y * blockDim.y + threadIdx.y;
    if (I >= vc.n[0] || J >= vc.n[1]) return;
    const int g0 = vc.n[0] + 2, g1 = vc.n[1] + 2;
    const T* xgb = xg + (long long)c * g0 * g1 * (vc.n[2] + 2);
    // ghosted index of (I, J, K) is (I+1, J+1, K+1)
    const long long o = (long long)(I + 1) + (long long)g0 * ((long long)(J + 1) + (long long)g1 * (K + 1));
    T nb[3][3][3];
#pragma unroll
    for (int z = 0; z < 3; ++z)
#pragma unroll
        for (int y = 0; y < 3; ++y)
#pragma unroll
            for (int x = 0; x < 3; ++x) nb[z][y][x] = xgb[o + (x - 1) + (long long)g0 * ((y - 1) + (long long)g1 * (z - 1))];
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
            {
                const int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                if (i >= nf0 || j >= nf1 || k >= nf2) continue;
                const int sx = ii ? 1 : -1, sy = jj ? 1 : -1, sz = kk ? 1 : -1;
                T acc = T(0);
#pragma unroll
                for (int cc = 0; cc < 2; ++cc)
#pragma unroll
                    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                        for (int aa = 0; aa < 2; ++aa)
                        {
                            T w = aa ? T(0.25) : T(0.75);
                            w *= bb ? T(0.25) : T(0.75);
                            w *= cc ? T(0.25) : T(0.75);
                            acc += w * nb[1 + cc * sz][1 + bb * sy][1 + aa * sx];
                        }
                const long long fo = (long long)i + (long long)nf0 * ((long long)j + (long long)nf1 * k);
                xf[(long long)c * Nf + fo] += acc;
            }
}

} // namespace vd
#endif
