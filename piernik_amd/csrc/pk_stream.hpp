// pk_stream.hpp -- the fused RK2 sweep as a stream (included by
// pk_kernels.hip inside its namespace; uses plane_of, tile helpers, floors,
// the timestep fold and the periodic images defined there).
//
// One wave per workgroup walks a stream of positions along the sweep.  TP
// pencils (consecutive along the first transverse direction: x for y and z
// sweeps, y for x sweeps) advance together; each pencil contributes
// L = n + 8 positions (its interior and four cells on each side; for a
// block periodic on itself whose ghosts are left stale, the four on each
// side are read from the interior cells they image), and the pencils follow
// each other: the stream is
//     position s  ->  pencil group s / L, cell (s mod L) - 4 + g.
// Per iteration the wave takes C = 64 / TP new positions per pencil (lane
// (p, i): pencil p, position base + i) and finishes everything the new
// values complete, lagging behind the loads:
//     q(s)                  primitives of the loaded cell
//     faces of cell s-1     from q(s-2), q(s-1), q(s)
//     F1 at (s-2)+1/2       stage-1 flux
//     U*(s-2)               stage-1 state (floors), its primitives qs
//     faces* of cell s-3    from qs(s-4), qs(s-3), qs(s-2)
//     F2 at (s-4)+1/2       stage-2 flux
//     U1(s-4)               the new state (floors, timestep fold, stored)
// A lane reads its lagged neighbours from LDS: positions of this iteration
// from the shared buffer, those of the previous iteration from small carry
// areas that the last lanes fill before moving on.  Every cell's two RK2
// stages are computed once, with the operands and operations of the tile
// kernel in the same order (bit-identical results); only the 8 halo
// positions per pencil, and the start-up and trailing positions of a range
// that starts or ends inside a pencil group, are extra work.  A wave is a
// whole workgroup, so the barriers between phases only order LDS accesses
// within the wave.
//
// The stream's pencil groups are split among the W workgroups of a launch:
// workgroup w owns groups [w*G/W, (w+1)*G/W) of the G groups, so the ranges
// differ by at most one group.  A range that starts or ends inside a group
// (items of the list form) starts 8 positions early, so that its first
// owned outputs have their stencil, and loads 4 positions past its end; a
// range writes only outputs at owned positions.

// LDS carry areas: slot offsets (in positions) of each lagged quantity in
// the carry block, and how many positions it keeps.
struct StreamCarry {
    static constexpr int Q = 0, NQ = 2;    // primitives, lags 1 and 2
    static constexpr int FP = 2, NFP = 1;  // right face state, lag 1
    static constexpr int F1 = 3, NF1 = 1;  // stage-1 flux, lag 1
    static constexpr int QS = 4, NQS = 2;  // stage-1 primitives, lags 1 and 2
    static constexpr int GP = 6, NGP = 1;  // stage-2 right face state, lag 1
    static constexpr int F2 = 7, NF2 = 1;  // stage-2 flux, lag 1
    static constexpr int N = 8;            // positions kept in the carry block
    static constexpr int ULAG = 4;         // U0 is read 2 and 4 positions back
};

// Orders the LDS accesses of the lanes of the one wave that is the whole
// workgroup.  The LDS unit performs a wave's accesses in the order the wave
// issues them, so a write is seen by any later read of the same wave; only
// the compiler must keep them in program order, which a fence at wavefront
// scope does without waiting for their completion.
__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

// Stores of a fluid's new state through raw buffer resources, one per three
// planes (pk::stream_stores_fit(sv): each range, 24 sv bytes, below 2^31).
// Plane k is at scalar offset (k % 3) * 8 sv of resource k / 3; a lane whose
// vector offset is the range's size (drop()) is out of range with any of
// these scalar offsets, and the hardware drops its store.  Both offsets stay
// below 2^31, so their sum does not wrap.
template <int NV>
struct PlaneStores {
    __amdgpu_buffer_rsrc_t r[(NV + 2) / 3];
    unsigned plane_bytes;
    __device__ __forceinline__ PlaneStores(double *dst, long long sv) : plane_bytes((unsigned)sv * 8u)
    {
#pragma unroll
        for (int g = 0; g < (NV + 2) / 3; g++)
            // dword 3 of a gfx9 buffer resource: 32-bit data format, no swizzle
            r[g] = __builtin_amdgcn_make_buffer_rsrc(dst + 3ll * g * sv, (short)0, (int)(3u * plane_bytes), 0x00020000);
    }
    __device__ __forceinline__ unsigned drop() const { return 3u * plane_bytes; }
    __device__ __forceinline__ void store(int plane, unsigned voff, double v) const
    {
        typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r[plane / 3], voff,
                                              (int)((unsigned)(plane % 3) * plane_bytes), 0);
    }
};

// Slots per variable of the U0 buffer of a stream workgroup.
template <int TP>
__host__ __device__ constexpr int stream_su()
{
    return TP * (64 / TP + StreamCarry::ULAG);
}

// The phase buffer A and the carry block K share one LDS array: per
// variable, 64 slots of A then StreamCarry::N positions x TP of K.
template <int TP>
__host__ __device__ constexpr int stream_stride()
{
    return 64 + StreamCarry::N * TP;
}

// The positions [s0, s0 + span) of a block's stream, walked by the one wave
// of the calling workgroup.  AK (NV * stream_stride<TP>()) and U
// (NV * stream_su<TP>()) are the workgroup's LDS; the wave's earlier
// contents there are never read for an owned output.  slot_id picks the
// timestep slot.  IMG: 1 the launch writes periodic ghost images
// (a.images == 1), 0 it writes none (a.images 0 or 2), -1 as a.images says
// (lists, whose blocks differ); a fixed IMG keeps the image code, and the
// registers it holds, out of kernels that never run it.
template <int DIR, class P, int LIM, int TP, int IMG>
__device__ __forceinline__ void stream_range(const SweepArgs &a, const int s0, const int span, unsigned slot_id,
                                             double *AK, double *U)
{
    constexpr int NV = P::NV;
    constexpr int C = 64 / TP;
    constexpr int UL = StreamCarry::ULAG;
    constexpr int SU = stream_su<TP>();
    constexpr int S = stream_stride<TP>();  // per-variable stride of A and K
    static_assert(C >= UL, "stream shape");
    double *const A = AK, *const K = AK + 64;
    constexpr int ds = DIR;
    constexpr int d1 = (DIR == 0) ? 1 : 0;
    constexpr int d2 = (DIR == 2) ? 1 : 2;

    const int lane = threadIdx.x;
    const int p = lane % TP, i = lane / TP;
    const int L = a.n[ds] + 8;
    const int n1 = a.n[d1];
    const int npen = n1 * a.n[d2];
    const int ngroups = (npen + TP - 1) / TP;
    const int npos = ngroups * L;
    const int s1 = min(s0 + span, npos);
    // A range that starts at a pencil group's first position needs no early
    // start: the group's own 8 leading positions (its ghost cells) fill the
    // pipeline before its first output.  One that ends at a group's end (or
    // the stream's) needs no loads past it: the group's last output, 4
    // positions before its end, is finished when its last position is loaded.
    const int first = (s0 % L == 0) ? s0 : s0 - 8;
    const int end = (s1 % L == 0) ? s1 : s1 + 4;
    const long long st = (DIR == 0) ? 1ll : ((DIR == 1) ? (long long)a.sy : (long long)a.sz);

    // LDS slots read at each lag (loop-invariant): this iteration's positions
    // from A, earlier ones from the quantity's carry positions in K
    auto rd = [&](int lag, int koff, int klag) -> const double * {
        return (i >= lag) ? &A[TP * (i - lag) + p] : &K[TP * (koff + klag + i - lag) + p];
    };
    const double *q1 = rd(1, StreamCarry::Q, StreamCarry::NQ), *q2 = rd(2, StreamCarry::Q, StreamCarry::NQ);
    const double *fp1 = rd(1, StreamCarry::FP, StreamCarry::NFP);
    const double *f11 = rd(1, StreamCarry::F1, StreamCarry::NF1);
    const double *qs1 = rd(1, StreamCarry::QS, StreamCarry::NQS), *qs2 = rd(2, StreamCarry::QS, StreamCarry::NQS);
    const double *gp1 = rd(1, StreamCarry::GP, StreamCarry::NGP);
    const double *f21 = rd(1, StreamCarry::F2, StreamCarry::NF2);
    double *const aw = &A[TP * i + p];
    // carry slot this lane fills for a quantity keeping klag positions (null: none)
    auto cw = [&](int koff, int klag) -> double * {
        return (i >= C - klag) ? &K[TP * (koff + i - (C - klag)) + p] : nullptr;
    };
    double *const cq = cw(StreamCarry::Q, StreamCarry::NQ), *const cfp = cw(StreamCarry::FP, StreamCarry::NFP);
    double *const cf1 = cw(StreamCarry::F1, StreamCarry::NF1), *const cqs = cw(StreamCarry::QS, StreamCarry::NQS);
    double *const cgp = cw(StreamCarry::GP, StreamCarry::NGP), *const cf2 = cw(StreamCarry::F2, StreamCarry::NF2);
    double *const uw = &U[TP * (i + UL) + p];
    const double *const u2 = &U[TP * (i + UL - 2) + p], *const u4 = &U[TP * (i + UL - 4) + p];
    double *const uc = (i >= C - UL) ? &U[TP * (i + UL - C) + p] : nullptr;

    // position of this lane in the stream: cell j of pencil group grp, its
    // pencil's transverse coordinates, advanced by C per iteration
    // (a position before the stream, in the first workgroup's start-up, is
    // in group -1, whose pencils do not exist)
    int sp = first + i;
    int grp, j, c1, c2;
    if (sp < 0) {
        grp = -1;
        j = sp + L;
        c1 = p - TP;
        c2 = 0;
    } else {
        grp = sp / L;
        j = sp - grp * L;
        const int pf = grp * TP + p;
        c1 = pf % n1;
        c2 = pf / n1;
    }
    auto advance = [&]() {
        sp += C;
        j += C;
        while (j >= L) {
            j -= L;
            grp++;
            c1 += TP;
            while (c1 >= n1) {
                c1 -= n1;
                c2++;
            }
        }
    };
    // global index (plane 0) of cell j of this lane's pencil; lanes outside
    // the stream or past the last pencil read a valid cell of the block
    // (wrap: a ghost position reads the interior cell it images, for a block
    // periodic on itself whose ghosts this launch leaves unwritten)
    const bool wrap = a.images == 2;
    const int n = a.n[ds];
    auto cell_index = [&](int jj, bool &pv, bool wr) -> long long {
        pv = grp >= 0 && (grp * TP + p) < npen;
        const int cc1 = pv ? c1 : 0, cc2 = pv ? c2 : 0;
        if (wr) jj = (jj < 4) ? jj + n : ((jj >= n + 4) ? jj - n : jj);
        int xyz[3];
        xyz[ds] = a.g[ds] - 4 + jj;
        xyz[d1] = a.g[d1] + cc1;
        xyz[d2] = a.g[d2] + cc2;
        return (long long)xyz[0] + (long long)a.sy * xyz[1] + (long long)a.sz * xyz[2];
    };

    unsigned flags = 0u;
    const double huge = 1.7976931348623157e308;
    double mn[3] = {huge, huge, huge}, cmax = 0.0;
    const int g = a.g[ds];

    double W[NV];
    {
        bool pv;
        const long long idx = cell_index(j, pv, wrap);
#pragma unroll
        for (int k = 0; k < NV; k++) W[k] = a.src[(long long)plane_of<DIR>(k) * a.sv + idx];
    }
    // Wait for these loads before the loop: the compiler's wait at the top of
    // the loop then only has to cover the loads of the previous iteration,
    // which the stores issued after them do not delay.
    __builtin_amdgcn_s_waitcnt(0);
    // outputs lag 4 positions behind the loads: the last owned output,
    // position s1-1, is finished when position s1+3 (or the group's last
    // position) is loaded
    const int iters = (end - first + C - 1) / C;
    for (int it = 0; it < iters; it++) {
        const int spc = sp, jc = j, cc1 = c1, cc2 = c2;  // this iteration's position and cell
        bool pv;
        const long long idx = cell_index(jc, pv, false);
        advance();

        // U0 and primitives of position sp
        double q[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) uw[v * SU] = W[v];
        prim<P>(W, a.gam_1, q);
#pragma unroll
        for (int v = 0; v < NV; v++) aw[v * S] = q[v];
        wave_lds_order();

        // faces of cell sp-1 from q(sp-2), q(sp-1), q(sp)
        double fp[NV], fm[NV];
        {
            double qa[NV], qb[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                qa[v] = q2[v * S];
                qb[v] = q1[v * S];
            }
            cell_faces<P, LIM>(a, qa, qb, q, fp, fm);
        }
        if (cq) {
#pragma unroll
            for (int v = 0; v < NV; v++) cq[v * S] = q[v];
        }
        wave_lds_order();
#pragma unroll
        for (int v = 0; v < NV; v++) aw[v * S] = fp[v];
        wave_lds_order();

        // stage-1 flux at (sp-2)+1/2: right face of sp-2, left face of sp-1;
        // the carry of fp is written once the lagged faces are read, so fp
        // holds no registers through the solver
        double F1[NV];
        {
            double fl[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) fl[v] = fp1[v * S];
            wave_lds_order();
            if (cfp) {
#pragma unroll
                for (int v = 0; v < NV; v++) cfp[v * S] = fp[v];
            }
            riemann<P>(a, fl, fm, F1);
        }
        wave_lds_order();
#pragma unroll
        for (int v = 0; v < NV; v++) aw[v * S] = F1[v];
        wave_lds_order();

        // stage-1 state of cell sp-2 and its primitives
        double qs[NV];
        {
            double Fm[NV], W0[NV], Ws[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                Fm[v] = f11[v * S];
                W0[v] = u2[v * SU];
            }
            cell_update<P>(W0, a.c1, Fm, F1, Ws);
            unsigned fl = 0u;
            double dm = 0.0;
            floors<P>(Ws, a.smalld, a.smallei, a.use_smalld, a.use_smallei, a.ei_flag, fl, dm);
            // cells g-2 .. g+n+1 (j-2 in [2, L-2)) raise flags, the interior counts mass
            const int jj = jc - 2;
            const bool own = pv && spc - 2 >= s0 && spc - 2 < s1 && jc >= 2;  // same pencil as spc
            if (own && jj >= 2 && jj < L - 2) flags |= fl;
            pk::mass_add(a.dmass, (own && jj >= 4 && jj < L - 4) ? dm : 0.0);
            prim<P>(Ws, a.gam_1, qs);
        }
        if (cf1) {
#pragma unroll
            for (int v = 0; v < NV; v++) cf1[v * S] = F1[v];
        }
        wave_lds_order();
#pragma unroll
        for (int v = 0; v < NV; v++) aw[v * S] = qs[v];
        wave_lds_order();

        // stage-2 faces of cell sp-3 from qs(sp-4), qs(sp-3), qs(sp-2)
        double gp[NV], gm[NV];
        {
            double qa[NV], qb[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                qa[v] = qs2[v * S];
                qb[v] = qs1[v * S];
            }
            cell_faces<P, LIM>(a, qa, qb, qs, gp, gm);
        }
        if (cqs) {
#pragma unroll
            for (int v = 0; v < NV; v++) cqs[v * S] = qs[v];
        }
        wave_lds_order();
#pragma unroll
        for (int v = 0; v < NV; v++) aw[v * S] = gp[v];
        wave_lds_order();

        // stage-2 flux at (sp-4)+1/2 (the carry of gp as that of fp)
        double F2[NV];
        {
            double fl[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) fl[v] = gp1[v * S];
            wave_lds_order();
            if (cgp) {
#pragma unroll
                for (int v = 0; v < NV; v++) cgp[v * S] = gp[v];
            }
            riemann<P>(a, fl, gm, F2);
        }
        wave_lds_order();
#pragma unroll
        for (int v = 0; v < NV; v++) aw[v * S] = F2[v];
        wave_lds_order();

        // U0 of the next iteration's positions, loaded before this
        // iteration's stores.  Vector memory counts loads and stores in one
        // in-order counter; as the stores below are issued on every path,
        // the wait for these loads at the top of the next iteration leaves
        // the stores in flight.
        {
            bool pv2;
            const long long idx2 = cell_index(j, pv2, wrap);
#pragma unroll
            for (int k = 0; k < NV; k++) W[k] = a.src[(long long)plane_of<DIR>(k) * a.sv + idx2];
        }

        // the new state of cell sp-4
        {
            double Fm[NV], W0[NV], Wn[NV];
#pragma unroll
            for (int v = 0; v < NV; v++) {
                Fm[v] = f21[v * S];
                W0[v] = u4[v * SU];
            }
            cell_update<P>(W0, a.c2, Fm, F2, Wn);
            unsigned fl = 0u;
            double dm = 0.0;
            floors<P>(Wn, a.smalld, a.smallei, a.use_smalld, a.use_smallei, a.ei_flag, fl, dm);
            const int jo = jc - 4;  // output cell: interior when jo in [4, L-4)
            const bool out = pv && spc - 4 >= s0 && spc - 4 < s1 && jo >= 4 && jo < L - 4;
            if (out) flags |= fl;
            pk::mass_add(a.dmass, out ? dm : 0.0);
            if constexpr (P::MAG) Wn[8] = Wn[8] * a.psi_fac;
            const long long oidx = idx - 4 * st;
            if (out) {
                if (a.dtslots) dt_cell<DIR, P>(a, Wn, mn, cmax);
                if ((IMG < 0) ? (a.images == 1) : (IMG == 1)) {
                    int xyz[3];
                    xyz[ds] = g - 4 + jo;
                    xyz[d1] = a.g[d1] + cc1;
                    xyz[d2] = a.g[d2] + cc2;
                    write_images<DIR, P>(a, xyz, oidx, Wn);
                }
            }
            // every lane issues the stores: a lane without an output stores
            // past the end of the plane's buffer range, which the hardware drops
            const PlaneStores<NV> ps(a.dst, a.sv);
            const unsigned voff = out ? (unsigned)oidx * 8u : ps.drop();
#pragma unroll
            for (int k = 0; k < NV; k++) ps.store(plane_of<DIR>(k), voff, Wn[k]);
        }
        if (cf2) {
#pragma unroll
            for (int v = 0; v < NV; v++) cf2[v * S] = F2[v];
        }
        // U0 of the last UL positions goes in front of the next iteration's
        if (uc) {
#pragma unroll
            for (int v = 0; v < NV; v++) uc[v * SU] = uw[v * SU];
        }
        wave_lds_order();
    }
    if (a.dtslots) dt_reduce(a, mn, cmax, slot_id);
    if (flags) atomicOr(a.flags, flags);
}

// Both stream kernels are held to two waves per SIMD (256 registers): at
// 253-256 VGPRs the allocator otherwise spills into AGPRs for some template
// instances and halves their occupancy.
template <int DIR, class P, int LIM, int TP, bool IMG>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_stream(const SweepArgs a,
                                                                                     const int ngroups)
{
    // the quantity of the current phase (slot TP*i + p) and the carry block
    __shared__ double AK[P::NV * stream_stride<TP>()];
    // U0 with ULAG positions of the previous iteration in front
    __shared__ double U[P::NV * stream_su<TP>()];
    // XCD-aware order: workgroup i runs on XCD i % 8, so each XCD gets a
    // contiguous range of the stream; neighbouring pencil groups of a y or z
    // sweep, which share cache lines, then share an L2
    int bid = blockIdx.x;
    const int nfull = (int)(gridDim.x / 8u) * 8;
    if (bid < nfull) bid = (bid % 8) * (nfull / 8) + bid / 8;
    const int L = a.n[DIR] + 8;
    const int g0 = (int)((long long)bid * ngroups / gridDim.x), g1 = (int)((long long)(bid + 1) * ngroups / gridDim.x);
    stream_range<DIR, P, LIM, TP, IMG ? 1 : 0>(a, g0 * L, (g1 - g0) * L, blockIdx.x, AK, U);
}

// The stream over a block list: one-wave workgroups take items from a
// device queue until it is empty; item `item` of block items[item] is the
// range of a.out positions starting at (item - first[block]) * a.out of
// that block's stream.  Each wave leaves after its first fetch past the
// end, so the grid drains after at most nitems + gridDim.x fetches.
template <int DIR, class P, int LIM, int TP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_stream_list(const ListArgs L)
{
    __shared__ double AK[P::NV * stream_stride<TP>()];
    __shared__ double U[P::NV * stream_su<TP>()];
    for (;;) {
        unsigned got = 0u;
        if (threadIdx.x == 0) got = atomicAdd(L.queue, 1u);
        const unsigned item = (unsigned)__builtin_amdgcn_readfirstlane((int)__shfl((int)got, 0, 64));
        if (item >= L.nitems) break;
        SweepArgs a;
        const unsigned first = load_list_args<DIR>(L, item, a);
        stream_range<DIR, P, LIM, TP, -1>(a, (int)(item - first) * a.out, a.out, item, AK, U);
    }
}
