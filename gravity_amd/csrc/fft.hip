// P3M mesh: deposit, 2D FFT, Green's-function multiply, inverse 2D FFT
// (S/clear_grid.comp, S/build_grid.comp:53-69, S/fft_init.comp,
// S/fft_shared.comp, S/convolve.comp; K/GravityPipeline.kt:907-973).
//
// Three launches per force build:
//   rows forward : per mesh row, gather the fixed-point CIC deposit of the
//                  bodies of the three physical rows that can reach it
//                  (integer LDS atomics: exact and order-free), convert to
//                  float (S/fft_init.comp) and transform along x;
//   columns      : per block of columns, forward transform along y, multiply
//                  by the precomputed spectrum G(k) = deconvolution / (W H) *
//                  (Kx(k) + i Ky(k)) / M, inverse transform along y;
//   rows inverse : per needed mesh row, inverse transform along x. Because
//                  the x and y force kernels are real, one complex inverse
//                  yields Fx in the real part and Fy in the imaginary part.
// Transforms are Stockham autosort in LDS: one lane owns one radix-8
// butterfly of a stage (radix 4 or 2 for the last stage, a few butterflies a
// lane), so a lane holds 8 complex values and many waves share a CU.
// Twiddles come from an FP64-rounded table of 2048 entries, staged in LDS by
// every workgroup. Under walls the mesh doubles;
// rows that hold only zero padding are neither loaded nor transformed in
// the forward row pass, and only rows the force gather reads are produced by
// the inverse passes.
#include "common.hpp"
#include "kernels.hpp"

namespace gs {

// A mesh line holds at most 2048 values. Row workgroups take one 2048-value
// row (or several shorter ones) with a radix-8 butterfly per lane and stage;
// column workgroups take 4096 values: 2 to 4 workgroups share a CU, so one's
// global loads overlap another's LDS stages.
constexpr int ROW_LDS = 2048;   // complex values per row workgroup
constexpr int ROW_NT = 256;     // lanes per row workgroup
constexpr int COL_LDS = 4096;   // complex values per column workgroup
constexpr int COL_NT = 512;     // lanes per column workgroup

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// Multiply by exp(DIR i pi / 2): -i forward, +i inverse.
template <int DIR>
__device__ __forceinline__ float2 mul_w4(float2 a) {
    return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int DIR>
__device__ __forceinline__ void dft2(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3) {
    float2 a0 = cadd(v0, v2), a1 = csub(v0, v2);
    float2 a2 = cadd(v1, v3), a3 = mul_w4<DIR>(csub(v1, v3));
    v0 = cadd(a0, a2);
    v2 = csub(a0, a2);
    v1 = cadd(a1, a3);
    v3 = csub(a1, a3);
}

template <int DIR>
__device__ __forceinline__ void dft8(float2* v) {
    const float c = 0.70710678118654752440f;
    float2 a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        a[k] = cadd(v[k], v[k + 4]);
        b[k] = csub(v[k], v[k + 4]);
    }
    b[1] = cmul(b[1], make_float2(c, DIR * c));
    b[2] = mul_w4<DIR>(b[2]);
    b[3] = cmul(b[3], make_float2(-c, DIR * c));
    dft4<DIR>(a[0], a[1], a[2], a[3]);
    dft4<DIR>(b[0], b[1], b[2], b[3]);
    v[0] = a[0]; v[2] = a[1]; v[4] = a[2]; v[6] = a[3];
    v[1] = b[0]; v[3] = b[1]; v[5] = b[2]; v[7] = b[3];
}

// Radix 16 as 4 x 4: four 4-point DFTs over the inputs 4 apart, the
// twiddles W16^(n2 k1), four 4-point DFTs across them; outputs in order.
template <int DIR>
__device__ __forceinline__ void dft16(float2* v) {
    const float c1 = 0.92387953251128675613f;  // cos(pi / 8)
    const float s1 = 0.38268343236508977173f;  // sin(pi / 8)
    const float c2 = 0.70710678118654752440f;  // cos(pi / 4)
    float2 y[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
        float2 a0 = v[n2], a1 = v[4 + n2], a2 = v[8 + n2], a3 = v[12 + n2];
        dft4<DIR>(a0, a1, a2, a3);
        y[4 * n2 + 0] = a0;
        y[4 * n2 + 1] = a1;
        y[4 * n2 + 2] = a2;
        y[4 * n2 + 3] = a3;
    }
    // W16^m = (cos(2 pi m / 16), DIR sin(2 pi m / 16)) for m = n2 k1
    y[5] = cmul(y[5], make_float2(c1, DIR * s1));     // m = 1
    y[6] = cmul(y[6], make_float2(c2, DIR * c2));     // m = 2
    y[7] = cmul(y[7], make_float2(s1, DIR * c1));     // m = 3
    y[9] = cmul(y[9], make_float2(c2, DIR * c2));     // m = 2
    y[10] = mul_w4<DIR>(y[10]);                       // m = 4
    y[11] = cmul(y[11], make_float2(-c2, DIR * c2));  // m = 6
    y[13] = cmul(y[13], make_float2(s1, DIR * c1));   // m = 3
    y[14] = cmul(y[14], make_float2(-c2, DIR * c2));  // m = 6
    y[15] = cmul(y[15], make_float2(-c1, -DIR * s1)); // m = 9
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        float2 b0 = y[k1], b1 = y[4 + k1], b2 = y[8 + k1], b3 = y[12 + k1];
        dft4<DIR>(b0, b1, b2, b3);
        v[k1] = b0;
        v[k1 + 4] = b1;
        v[k1 + 8] = b2;
        v[k1 + 12] = b3;
    }
}

// One Stockham stage of radix R on nl lines of length L held in LDS;
// element k of line l lives at buf[l * ls + k * es]. With NT lanes and at
// most E values in LDS a lane takes Q = E / (R NT) butterflies. jFast:
// consecutive lanes take consecutive butterflies of one line (rows);
// otherwise the same butterfly of consecutive lines (columns).
template <int DIR, int R, int Q, int NT>
__device__ __forceinline__ void stockham_stage(float2* buf, int L, int nl, int ls, int es, int Ns,
                                               bool jFast, const float2* __restrict__ tw,
                                               int twLen) {
    const int perLine = L / R;
    const int total = perLine * nl;
    float2 v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        int bi = (int)threadIdx.x + q * NT;
        if (bi < total) {
            int line = jFast ? bi / perLine : bi % nl, j = jFast ? bi % perLine : bi / nl;
#pragma unroll
            for (int r = 0; r < R; r++) v[q][r] = buf[line * ls + (j + r * perLine) * es];
        }
    }
    __syncthreads();
    const int twStep = twLen / (Ns * R);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        int bi = (int)threadIdx.x + q * NT;
        if (bi < total) {
            int line = jFast ? bi / perLine : bi % nl, j = jFast ? bi % perLine : bi / nl;
            int k = j & (Ns - 1);
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    float2 w = tw[r * k * twStep];
                    if (DIR > 0) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            if (R == 16) dft16<DIR>(v[q]);
            else if (R == 8) dft8<DIR>(v[q]);
            else if (R == 4) dft4<DIR>(v[q][0], v[q][1], v[q][2], v[q][3]);
            else dft2<DIR>(v[q]);
            int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) buf[line * ls + (base + r * Ns) * es] = v[q][r];
        }
    }
    __syncthreads();
}

// Transforms nl lines (nl * L <= E). DIR = -1 forward, +1 inverse
// (unscaled, like S/fft_shared.comp). Stages of radix MAXR (8 or 16) first:
// the column transforms, whose workgroups hold four lines, take radix 16
// (one stage fewer per direction); the rows keep radix 8, which gives each of
// their lanes a butterfly per stage.
template <int DIR, int NT, int E, int MAXR = 8>
__device__ __forceinline__ void lds_fft(float2* buf, int L, int nl, int ls, int es, const float2* __restrict__ tw,
                        int twLen) {
    const bool jFast = es == 1;
    int Ns = 1;
    while (Ns < L) {
        int rem = L / Ns;
        if (MAXR >= 16 && rem >= 16) {
            stockham_stage<DIR, 16, (E / (16 * NT) > 0 ? E / (16 * NT) : 1), NT>(buf, L, nl, ls, es, Ns, jFast, tw, twLen);
            Ns *= 16;
        } else if (rem >= 8) {
            stockham_stage<DIR, 8, (E / (8 * NT) > 0 ? E / (8 * NT) : 1), NT>(buf, L, nl, ls, es, Ns, jFast, tw, twLen);
            Ns *= 8;
        } else if (rem == 4) {
            stockham_stage<DIR, 4, (E / (4 * NT) > 0 ? E / (4 * NT) : 1), NT>(buf, L, nl, ls, es, Ns, jFast, tw, twLen);
            Ns *= 4;
        } else {
            stockham_stage<DIR, 2, (E / (2 * NT) > 0 ? E / (2 * NT) : 1), NT>(buf, L, nl, ls, es, Ns, jFast, tw, twLen);
            Ns *= 2;
        }
    }
}

// The twiddle table (at most TW_LDS entries) staged in LDS with the first
// loads of a kernel: the stages then read it without a memory round trip.
// A lane's TW_LDS / NT loads are issued together (clamped to the table), then
// stored: one round trip, where a load-store loop waits for every load.
// A kernel loads its lane's part of the table first and stores it to LDS
// after issuing its other global loads, so those are in flight together.
constexpr int TW_LDS = 2048;
template <int NT>
struct Twiddles {
    static constexpr int PER = TW_LDS / NT;
    float2 v[PER];
    __device__ __forceinline__ void load(const float2* __restrict__ tw, int twLen) {
#pragma unroll
        for (int q = 0; q < PER; q++) v[q] = tw[min((int)threadIdx.x + q * NT, twLen - 1)];
    }
    __device__ __forceinline__ void store(float2* twl, int twLen) const {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int i = (int)threadIdx.x + q * NT;
            if (i < twLen) twl[i] = v[q];
        }
    }
};
template <int NT>
__device__ __forceinline__ void stage_twiddles(float2* twl, const float2* __restrict__ tw,
                                               int twLen) {
    Twiddles<NT> t;
    t.load(tw, twLen);
    t.store(twl, twLen);
}

// Mesh row of forward/inverse work row a: all rows on the torus; under walls
// rows 0..ph and the last one (cells at -1 wrap there).
__device__ __forceinline__ int mesh_row(const SimDesc& s, int a) {
    if (!s.bounded) return a;
    return a <= s.ph ? a : s.mh - 1;
}

// ---------------------------------------------------------------------------
// DENSITY: also store the fixed-point deposit of the item's rows to density
// (the readback of the density mesh, S/build_grid.comp's buffer).
// RUNS: a lane sums the fixed-point corner weights of its consecutive bodies
// in registers while they share a 3x3 window of mesh nodes (the nodes a
// physical cell's bodies reach) and adds each non-zero sum to LDS when a body
// leaves the window: in a crowded cell hundreds of bodies then cost a few LDS
// atomics instead of four each, on a handful of contended addresses. The
// corner weights round per body as before and integer sums do not depend on
// their order, so the deposit is the same either way.
template <bool DENSITY, bool RUNS>
__global__ void __launch_bounds__(ROW_NT) k_fft_rows_forward(const SimDesc* __restrict__ sims,
                                                             const LineWork* __restrict__ work,
                                                             const uint32_t* __restrict__ cellPrefix,
                                                             const float4* __restrict__ cellBody,
                                                             float2* __restrict__ spec,
                                                             const float2* __restrict__ tw, int twLen,
                                                             uint32_t* __restrict__ density) {
    __shared__ float2 data[ROW_LDS];
    __shared__ uint32_t mass[ROW_LDS];
    __shared__ float2 twl[TW_LDS];
    LineWork w = work[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[w.sim];
    const int L = s.mw;
    Twiddles<ROW_NT> tws;
    tws.load(tw, twLen);  // stored after the deposit's loads
    for (int i = (int)threadIdx.x; i < w.count * L; i += ROW_NT) mass[i] = 0u;
    __syncthreads();
    const float scale = s.massScale;
    const int first = w.first, count = w.count;
    // A body of physical row r deposits into mesh rows r - 1 .. r + 1 (under
    // walls a body of row 0 below the edge reaches the last mesh row, the
    // item's line ph + 1). The physical rows that reach the item's lines are
    // at most two ranges: rows [ra[q], rb[q]].
    int ra[2] = {0, 0}, rb[2] = {-1, -1}, nr = 0;
    if (count == 1) {
        const int y = mesh_row(s, first);
        if (s.bounded) {
            if (y == s.mh - 1) { ra[0] = 0; rb[0] = 0; }
            else { ra[0] = max(y - 1, 0); rb[0] = min(y + 1, s.ph - 1); }
            nr = 1;
        } else if (s.ph < 3) {
            ra[0] = 0; rb[0] = s.ph - 1; nr = 1;
        } else if (y == 0) {
            ra[0] = s.ph - 1; rb[0] = s.ph - 1; ra[1] = 0; rb[1] = 1; nr = 2;
        } else if (y == s.ph - 1) {
            ra[0] = y - 1; rb[0] = s.ph - 1; ra[1] = 0; rb[1] = 0; nr = 2;
        } else {
            ra[0] = y - 1; rb[0] = y + 1; nr = 1;
        }
    } else if (!s.bounded) {
        if (count + 2 >= s.ph) {
            ra[0] = 0; rb[0] = s.ph - 1; nr = 1;
        } else if (first == 0) {
            ra[0] = s.ph - 1; rb[0] = s.ph - 1; ra[1] = 0; rb[1] = count; nr = 2;
        } else if (first + count == s.ph) {
            ra[0] = first - 1; rb[0] = s.ph - 1; ra[1] = 0; rb[1] = 0; nr = 2;
        } else {
            ra[0] = first - 1; rb[0] = first + count; nr = 1;
        }
    } else {
        if (first <= s.ph) {
            ra[0] = max(first - 1, 0);
            rb[0] = min(min(first + count - 1, s.ph) + 1, s.ph - 1);
            nr = 1;
        }
        if (first + count - 1 >= s.ph + 1 && (nr == 0 || ra[0] > 0)) {
            ra[nr] = 0; rb[nr] = 0; nr++;
        }
    }
    // An item of one line (every row of a 2048-wide mesh) keeps the corners
    // in its mesh row; an item of several lines routes each corner to the
    // line of its mesh row, so each body of the rows is visited once.
    const int line0 = mesh_row(s, first);
    // RUNS: the open window's sums (row-major, rows ky.., columns kx..).
    uint32_t run[9];
#pragma unroll
    for (int k = 0; k < 9; k++) run[k] = 0u;
    int kx = -0x40000000, ky = -0x40000000;
    auto flush = [&]() {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int my = wrap_axis(ky + j, s.mh);
            int l1, l2;
            if (count == 1) {
                l1 = my == line0 ? 0 : -1;
                l2 = -1;
            } else {
                l1 = !s.bounded || my <= s.ph ? my - first : -1;
                l2 = s.bounded && my == s.mh - 1 ? s.ph + 1 - first : -1;
            }
            const bool in1 = (unsigned)l1 < (unsigned)count;
            const bool in2 = (unsigned)l2 < (unsigned)count;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const uint32_t v = run[j * 3 + i];
                run[j * 3 + i] = 0u;
                if (v == 0u || (!in1 && !in2)) continue;
                const int mx = wrap_axis(kx + i, s.mw);
                if (in1) atomicAdd(&mass[l1 * L + mx], v);
                if (in2) atomicAdd(&mass[l2 * L + mx], v);
            }
        }
    };
    // The rows of a range are contiguous in the cell segments.
    for (int q = 0; q < nr; q++) {
        {
            const int t0 = (int)cellPrefix[s.physBase + ra[q] * s.pw];
            const int t1 = (int)cellPrefix[s.physBase + (rb[q] + 1) * s.pw];
            // Eight bodies a lane per round, their loads issued together and
            // unconditional (clamped to the range).
            constexpr int UNROLL = 8;
            for (int t = t0 + (int)threadIdx.x; t < t1; t += UNROLL * ROW_NT) {
                float4 bds[UNROLL];
#pragma unroll
                for (int u = 0; u < UNROLL; u++) bds[u] = cellBody[min(t + u * ROW_NT, t1 - 1)];
#pragma unroll
                for (int u = 0; u < UNROLL; u++) {
                    if (t + u * ROW_NT >= t1) continue;
                    float4 bd = bds[u];
                    // S/build_grid.comp:53-69, round() half to even.
                    float gx = bd.x * (float)s.pw - 0.5f, gy = bd.y * (float)s.ph - 0.5f;
                    float flx = floorf(gx), fly = floorf(gy);
                    int bx = (int)flx, by = (int)fly;
                    float fx = gx - flx, fy = gy - fly;
                    float wgt[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy),
                                    (1.0f - fx) * fy, fx * fy};
                    if (RUNS) {
                        // The window of the body's physical cell: its corners
                        // are columns bx, bx + 1 of columns cx - 1 .. cx + 1.
                        const int wx = bx - (fx < 0.5f ? 1 : 0);
                        const int wy = by - (fy < 0.5f ? 1 : 0);
                        if (wx != kx || wy != ky) {
                            flush();
                            kx = wx;
                            ky = wy;
                        }
                        const int ox = bx - kx, oy = by - ky;
                        uint32_t qv[4];
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            float wk = wgt[k] * bd.z;
                            qv[k] = (uint32_t)rintf(wk * scale);
                        }
#pragma unroll
                        for (int j = 0; j < 3; j++) {
                            const uint32_t qa = j == oy ? qv[0] : (j == oy + 1 ? qv[2] : 0u);
                            const uint32_t qb = j == oy ? qv[1] : (j == oy + 1 ? qv[3] : 0u);
#pragma unroll
                            for (int i = 0; i < 3; i++)
                                run[j * 3 + i] += i == ox ? qa : (i == ox + 1 ? qb : 0u);
                        }
                        continue;
                    }
                    if (count == 1) {
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            int my = wrap_axis(by + (k >> 1), s.mh);
                            if (my != line0) continue;
                            int mx = wrap_axis(bx + (k & 1), s.mw);
                            float wk = wgt[k] * bd.z;
                            uint32_t qv = (uint32_t)rintf(wk * scale);
                            atomicAdd(&mass[mx], qv);
                        }
                        continue;
                    }
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int my = wrap_axis(by + (k >> 1), s.mh);
                        const int mx = wrap_axis(bx + (k & 1), s.mw);
                        // the item's line of mesh row my, and under walls the
                        // last line for the last mesh row
                        const int l1 = !s.bounded || my <= s.ph ? my - first : -1;
                        const int l2 = s.bounded && my == s.mh - 1 ? s.ph + 1 - first : -1;
                        const bool in1 = (unsigned)l1 < (unsigned)count;
                        const bool in2 = (unsigned)l2 < (unsigned)count;
                        if (!in1 && !in2) continue;
                        float wk = wgt[k] * bd.z;
                        uint32_t qv = (uint32_t)rintf(wk * scale);
                        if (in1) atomicAdd(&mass[l1 * L + mx], qv);
                        if (in2) atomicAdd(&mass[l2 * L + mx], qv);
                    }
                }
            }
        }
    }
    if (RUNS) flush();
    tws.store(twl, twLen);
    __syncthreads();
    if (DENSITY) {
        for (int i = (int)threadIdx.x; i < w.count * L; i += ROW_NT) {
            int l = i / L, x = i - l * L;
            density[s.meshBase + mesh_row(s, w.first + l) * s.mw + x] = mass[i];
        }
    }
    for (int i = (int)threadIdx.x; i < w.count * L; i += ROW_NT)
        data[i] = make_float2((float)mass[i] * s.invMassScale, 0.0f);
    __syncthreads();
    lds_fft<-1, ROW_NT, ROW_LDS>(data, L, w.count, L, 1, twl, twLen);
    for (int i = (int)threadIdx.x; i < w.count * L; i += ROW_NT) {
        int l = i / L, x = i - l * L;
        int y = mesh_row(s, w.first + l);
        spec[s.meshBase + y * s.mw + x] = data[i];
    }
}

// FORWARD_ONLY: stop after the forward transform and store every row (the
// density spectrum of the reference's validateFft, K/GravityPipeline.kt:378-419).
template <bool FORWARD_ONLY>
__global__ void __launch_bounds__(COL_NT) k_fft_columns(const SimDesc* __restrict__ sims,
                                                        const LineWork* __restrict__ work,
                                                        float2* __restrict__ spec,
                                                        const float2* __restrict__ green,
                                                        const float2* __restrict__ tw, int twLen) {
    __shared__ float2 data[COL_LDS];
    __shared__ float2 twl[TW_LDS];
    LineWork w = work[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[w.sim];
    const int H = s.mh, CB = w.count, c0 = w.first;
    Twiddles<COL_NT> tws;
    tws.load(tw, twLen);  // stored after the values' loads
    const float2* src = spec + s.meshBase;
    // A lane's COL_LDS / COL_NT values load together (a dead row or a lane
    // past the item reads its column's row 0), then go to LDS.
    constexpr int PER = COL_LDS / COL_NT;
    {
        float2 v[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int i = min((int)threadIdx.x + q * COL_NT, CB * H - 1);
            const int y = i / CB, c = i - y * CB;
            const bool live = !s.bounded || y <= s.ph || y == H - 1;
            v[q] = src[(live ? y * s.mw : 0) + c0 + c];
        }
        tws.store(twl, twLen);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int i = (int)threadIdx.x + q * COL_NT;
            if (i < CB * H) {
                const int y = i / CB;
                const bool live = !s.bounded || y <= s.ph || y == H - 1;
                data[i] = live ? v[q] : make_float2(0.0f, 0.0f);
            }
        }
    }
    // The spectrum values of this workgroup's nodes load now and arrive
    // during the forward transform (COL_LDS / COL_NT of them per lane).
    const float2* G = green + s.meshBase;
    float2 gv[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        int i = (int)threadIdx.x + q * COL_NT;
        if (i < CB * H) {
            int y = i / CB, c = i - y * CB;
            gv[q] = G[y * s.mw + c0 + c];
        }
    }
    __syncthreads();
    lds_fft<-1, COL_NT, COL_LDS, 16>(data, H, CB, 1, CB, twl, twLen);
    if (FORWARD_ONLY) {
        float2* out = spec + s.meshBase;
        for (int i = (int)threadIdx.x; i < CB * H; i += COL_NT) {
            int y = i / CB, c = i - y * CB;
            out[y * s.mw + c0 + c] = data[i];
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < PER; q++) {
        int i = (int)threadIdx.x + q * COL_NT;
        if (i < CB * H) {
            float2 z = cmul(data[i], gv[q]);
            data[i] = make_float2(z.x * s.invTotalMass, z.y * s.invTotalMass);
        }
    }
    __syncthreads();
    lds_fft<1, COL_NT, COL_LDS, 16>(data, H, CB, 1, CB, twl, twLen);
    float2* dst = spec + s.meshBase;
    for (int i = (int)threadIdx.x; i < CB * H; i += COL_NT) {
        int y = i / CB, c = i - y * CB;
        bool live = !s.bounded || y <= s.ph || y == H - 1;
        if (live) dst[y * s.mw + c0 + c] = data[i];
    }
}

__global__ void __launch_bounds__(ROW_NT) k_fft_rows_inverse(const SimDesc* __restrict__ sims,
                                                             const LineWork* __restrict__ work,
                                                             float2* __restrict__ spec,
                                                             const float2* __restrict__ tw, int twLen) {
    __shared__ float2 data[ROW_LDS];
    __shared__ float2 twl[TW_LDS];
    LineWork w = work[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[w.sim];
    const int L = s.mw;
    Twiddles<ROW_NT> tws;
    tws.load(tw, twLen);  // stored after the values' loads
    {
        // A lane's ROW_LDS / ROW_NT values load together, then go to LDS.
        constexpr int PER = ROW_LDS / ROW_NT;
        const int n = w.count * L;
        float2 v[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int i = min((int)threadIdx.x + q * ROW_NT, n - 1);
            const int l = i / L, x = i - l * L;
            v[q] = spec[s.meshBase + mesh_row(s, w.first + l) * s.mw + x];
        }
        tws.store(twl, twLen);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int i = (int)threadIdx.x + q * ROW_NT;
            if (i < n) data[i] = v[q];
        }
    }
    __syncthreads();
    lds_fft<1, ROW_NT, ROW_LDS>(data, L, w.count, L, 1, twl, twLen);
    for (int i = (int)threadIdx.x; i < w.count * L; i += ROW_NT) {
        int l = i / L, x = i - l * L;
        spec[s.meshBase + mesh_row(s, w.first + l) * s.mw + x] = data[i];
    }
}

// Adjoint of the force gather (the transposed CIC interpolation): every body
// spreads its float2 value over its four CIC nodes (no mass), with float LDS
// atomics, then the rows are transformed forward as in k_fft_rows_forward.
__global__ void __launch_bounds__(ROW_NT) k_fft_rows_forward_values(
    const SimDesc* __restrict__ sims, const LineWork* __restrict__ work,
    const uint32_t* __restrict__ cellPrefix, const float4* __restrict__ cellBody,
    const float2* __restrict__ value, float2* __restrict__ spec, const float2* __restrict__ tw,
    int twLen) {
    __shared__ float2 data[ROW_LDS];
    __shared__ float2 twl[TW_LDS];
    LineWork w = work[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[w.sim];
    const int L = s.mw;
    stage_twiddles<ROW_NT>(twl, tw, twLen);
    for (int i = (int)threadIdx.x; i < w.count * L; i += ROW_NT) data[i] = make_float2(0.0f, 0.0f);
    __syncthreads();
    for (int l = 0; l < w.count; l++) {
        int y = mesh_row(s, w.first + l);
        int r0, r1;
        if (s.bounded) {
            if (y == s.mh - 1) { r0 = 0; r1 = 0; }
            else { r0 = y - 1 < 0 ? 0 : y - 1; r1 = y + 1 > s.ph - 1 ? s.ph - 1 : y + 1; }
        } else if (s.ph < 3) {
            r0 = 0; r1 = s.ph - 1;
        } else {
            r0 = y - 1; r1 = y + 1;
        }
        for (int rr = r0; rr <= r1; rr++) {
            int r = s.bounded || s.ph < 3 ? rr : wrap_axis(rr, s.ph);
            int c0 = s.physBase + r * s.pw;
            int t0 = (int)cellPrefix[c0], t1 = (int)cellPrefix[c0 + s.pw];
            for (int t = t0 + (int)threadIdx.x; t < t1; t += ROW_NT) {
                float4 bd = cellBody[t];
                float2 v = value[__float_as_uint(bd.w)];
                float gx = bd.x * (float)s.pw - 0.5f, gy = bd.y * (float)s.ph - 0.5f;
                float flx = floorf(gx), fly = floorf(gy);
                int bx = (int)flx, by = (int)fly;
                float fx = gx - flx, fy = gy - fly;
                float wgt[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy,
                                fx * fy};
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    int my = wrap_axis(by + (k >> 1), s.mh);
                    if (my != y) continue;
                    int mx = wrap_axis(bx + (k & 1), s.mw);
                    atomicAdd(&data[l * L + mx].x, wgt[k] * v.x);
                    atomicAdd(&data[l * L + mx].y, wgt[k] * v.y);
                }
            }
        }
    }
    __syncthreads();
    lds_fft<-1, ROW_NT, ROW_LDS>(data, L, w.count, L, 1, twl, twLen);
    for (int i = (int)threadIdx.x; i < w.count * L; i += ROW_NT) {
        int l = i / L, x = i - l * L;
        int y = mesh_row(s, w.first + l);
        spec[s.meshBase + y * s.mw + x] = data[i];
    }
}

void launch_fft_rows_forward_values(hipStream_t st, const SimDesc* sims, const LineWork* work,
                                    int items, const uint32_t* cellPrefix, const float4* cellBody,
                                    const float2* value, float2* spec, const float2* twiddles,
                                    int twiddleLength) {
    if (items == 0) return;
    launch(k_fft_rows_forward_values, dim3(items), dim3(ROW_NT), 0, st, sims, work,
                       cellPrefix, cellBody, value, spec, twiddles, twiddleLength);
}

void launch_fft_rows_forward(hipStream_t st, const SimDesc* sims, const LineWork* work, int items,
                             const uint32_t* cellPrefix, const float4* cellBody, float2* spec,
                             const float2* twiddles, int twiddleLength, bool runs,
                             uint32_t* density) {
    if (items == 0) return;
    auto k = density ? (runs ? k_fft_rows_forward<true, true> : k_fft_rows_forward<true, false>)
                     : (runs ? k_fft_rows_forward<false, true> : k_fft_rows_forward<false, false>);
    launch(k, dim3(items), dim3(ROW_NT), 0, st, sims, work, cellPrefix, cellBody, spec, twiddles,
           twiddleLength, density);
}

void launch_fft_columns(hipStream_t st, const SimDesc* sims, const LineWork* work, int items,
                        float2* spec, const float2* green, const float2* twiddles,
                        int twiddleLength, bool forwardOnly) {
    if (items == 0) return;
    if (forwardOnly)
        launch(k_fft_columns<true>, dim3(items), dim3(COL_NT), 0, st, sims, work, spec, green,
               twiddles, twiddleLength);
    else
        launch(k_fft_columns<false>, dim3(items), dim3(COL_NT), 0, st, sims, work, spec, green,
               twiddles, twiddleLength);
}

void launch_fft_rows_inverse(hipStream_t st, const SimDesc* sims, const LineWork* work, int items,
                             float2* spec, const float2* twiddles, int twiddleLength) {
    if (items == 0) return;
    launch(k_fft_rows_inverse, dim3(items), dim3(ROW_NT), 0, st, sims, work, spec,
                       twiddles, twiddleLength);
}

int fft_row_lines_per_group(int L) { return L >= ROW_LDS ? 1 : ROW_LDS / L; }
int fft_columns_per_group(int H, int W) {
    int cb = COL_LDS / H;
    return cb < W ? cb : W;
}

}  // namespace gs
