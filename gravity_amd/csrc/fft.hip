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
// Transforms are Stockham autosort in LDS, specialised per line length (a
// power of two, 16..2048): a lane owns one radix-8 butterfly of a stage
// (radix 4 or 2 for the last stage, a few butterflies a lane), its twiddles
// sit in registers from host-built per-stage tables, and the stages
// ping-pong between two LDS buffers. Under
// walls the mesh doubles; rows that hold only zero padding are neither loaded
// nor transformed in the forward row pass, and only the rows and columns the
// force gather reads are produced by the inverse passes.
#include <stdexcept>

#include "common.hpp"
#include "green.hpp"
#include "kernels.hpp"

namespace gs {

// A mesh line holds at most 2048 values. Row workgroups take one 2048-value
// row (or several shorter ones) with a radix-8 butterfly per lane and stage;
// column workgroups take 4096 values: 2 workgroups share a CU, so one's
// global loads overlap another's LDS stages.
constexpr int ROW_LDS = 2048;   // complex values per row workgroup
constexpr int ROW_NT = 256;     // lanes per row workgroup
constexpr int ROWF_NT = 512;    // lanes per forward row workgroup (deposit and transform)
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

// Length of the FP64-rounded twiddle table the fixed tables are drawn from.
constexpr int TW_LDS = 2048;

// ---------------------------------------------------------------------------
// Fixed-length transforms. Mesh lines are powers of two from 16 to 2048, so a
// kernel switches once on its line length and runs a transform whose every
// index is a compile-time shift or mask:
//   - a lane's twiddles for every stage load once, at the kernel's start,
//     into registers (W_L^m = tw[m * TW_LDS / L] of the FP64-rounded table);
//     the inverse uses their conjugates, so one set serves both directions;
//   - the stages ping-pong between two LDS buffers: one barrier a stage;
//   - LDS positions are XOR-swizzled (bits 3..7 onto bits 0..4), which puts
//     the strided writes of the first stages (lane strides of 8 and 64
//     values) on distinct banks and keeps contiguous runs of 32 contiguous;
//   - complex products use fused multiply-adds (the transforms' results are
//     compared with a tolerance, never bit for bit).
__device__ __forceinline__ int fswz(int i) { return i ^ ((i >> 3) & 31); }

template <int DIR>
__device__ __forceinline__ float2 fmul_tw(float2 a, float2 w) {
    // a * w forward, a * conj(w) inverse
    if (DIR < 0) return make_float2(fmaf(a.x, w.x, -(a.y * w.y)), fmaf(a.x, w.y, a.y * w.x));
    return make_float2(fmaf(a.x, w.x, a.y * w.y), fmaf(a.y, w.x, -(a.x * w.y)));
}

// a * exp(DIR i pi / 4) = (1 + DIR i) a / sqrt(2); a * exp(3 DIR i pi / 4)
template <int DIR>
__device__ __forceinline__ float2 rot8_1(float2 a) {
    const float c = 0.70710678118654752440f;
    return DIR < 0 ? make_float2((a.x + a.y) * c, (a.y - a.x) * c)
                   : make_float2((a.x - a.y) * c, (a.y + a.x) * c);
}
template <int DIR>
__device__ __forceinline__ float2 rot8_3(float2 a) {
    const float c = 0.70710678118654752440f;
    return DIR < 0 ? make_float2((a.y - a.x) * c, -(a.x + a.y) * c)
                   : make_float2(-(a.x + a.y) * c, (a.x - a.y) * c);
}

template <int DIR>
__device__ __forceinline__ void fdft8(float2* v) {
    float2 a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        a[k] = cadd(v[k], v[k + 4]);
        b[k] = csub(v[k], v[k + 4]);
    }
    b[1] = rot8_1<DIR>(b[1]);
    b[2] = mul_w4<DIR>(b[2]);
    b[3] = rot8_3<DIR>(b[3]);
    dft4<DIR>(a[0], a[1], a[2], a[3]);
    dft4<DIR>(b[0], b[1], b[2], b[3]);
    v[0] = a[0]; v[2] = a[1]; v[4] = a[2]; v[6] = a[3];
    v[1] = b[0]; v[3] = b[1]; v[5] = b[2]; v[7] = b[3];
}

template <int R, int DIR>
__device__ __forceinline__ void fdft(float2* v) {
    if constexpr (R == 8) fdft8<DIR>(v);
    else if constexpr (R == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
    else dft2<DIR>(v);
}

// Radix of the stage that starts at span ns of a line of L values.
__host__ __device__ constexpr int fixed_radix_of(int L, int maxr, int ns) {
    return L / ns >= maxr ? maxr : (L / ns >= 8 ? 8 : L / ns);
}
template <int L, int MAXR>
__host__ __device__ constexpr int fixed_radix(int ns) { return fixed_radix_of(L, MAXR, ns); }

// The twiddle tables of the fixed-length transforms, one per line length L
// (16..2048, largest radix 8), each holding the twiddles W_L^(r k L / (ns R))
// of every stage after the first, laid out [stage][r - 1][k] (k < ns): the
// lanes of a wave load one r's twiddles from contiguous entries.
__host__ __device__ constexpr int fixed_table_size(int L, int maxr) {
    int n = 0;
    for (int ns = 1; ns < L; ns *= fixed_radix_of(L, maxr, ns))
        if (ns > 1) n += (fixed_radix_of(L, maxr, ns) - 1) * ns;
    return n;
}
__host__ __device__ constexpr int fixed_table_offset(int L, int maxr) {
    int off = 0;
    for (int l = 16; l <= 2048; l *= 2) {
        if (l == L && maxr == 8) return off;
        off += fixed_table_size(l, 8);
    }
    return off;
}
constexpr int FIXED_TABLE_ENTRIES = fixed_table_offset(0, 8);

// A WG's NL lines of L values: element e of line l at l * L + e (JFAST, rows:
// consecutive lanes take consecutive butterflies of a line) or at e * NL + l
// (columns: consecutive lanes take the same butterfly of consecutive lines).
template <int L, int NL, bool JFAST>
__device__ __forceinline__ int fixed_pos(int line, int e) {
    return fswz(JFAST ? line * L + e : e * NL + line);
}

template <int L, int NL, int NT, int MAXR, int NS = 1>
__host__ __device__ constexpr int fixed_twiddle_count() {
    if constexpr (NS >= L) {
        return 0;
    } else {
        constexpr int R = fixed_radix<L, MAXR>(NS);
        constexpr int Q = (NL * (L / R) + NT - 1) / NT;
        return (NS > 1 ? Q * (R - 1) : 0) + fixed_twiddle_count<L, NL, NT, MAXR, NS * R>();
    }
}

// A lane's twiddles of every stage, from the table of (L, MAXR); `table`
// points at that table's first stage.
template <int L, int NL, int NT, int MAXR, bool JFAST, int NS = 1>
__device__ __forceinline__ void fixed_load_twiddles(float2* w, const float2* __restrict__ table) {
    if constexpr (NS < L) {
        constexpr int R = fixed_radix<L, MAXR>(NS);
        constexpr int PER_LINE = L / R;
        constexpr int Q = (NL * PER_LINE + NT - 1) / NT;
        if constexpr (NS > 1) {
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int bi = (int)threadIdx.x + q * NT;
                const int j = JFAST ? bi % PER_LINE : bi / NL;
                const int k = j & (NS - 1);
#pragma unroll
                for (int r = 1; r < R; r++) w[q * (R - 1) + r - 1] = table[(r - 1) * NS + k];
            }
            fixed_load_twiddles<L, NL, NT, MAXR, JFAST, NS * R>(w + Q * (R - 1), table + (R - 1) * NS);
        } else {
            fixed_load_twiddles<L, NL, NT, MAXR, JFAST, NS * R>(w, table);
        }
    }
}

// The stages from span NS on: read src, write dst, barrier, swap. Returns the
// buffer that holds the transform (natural order). The caller has stored the
// input into src at fixed_pos and passed a barrier.
template <int DIR, int L, int NL, int NT, int MAXR, bool JFAST, int NS = 1>
__device__ __forceinline__ float2* fixed_stages(float2* src, float2* dst, const float2* w) {
    if constexpr (NS >= L) {
        return src;
    } else {
        constexpr int R = fixed_radix<L, MAXR>(NS);
        constexpr int PER_LINE = L / R;
        constexpr int TOTAL = NL * PER_LINE;
        constexpr int Q = (TOTAL + NT - 1) / NT;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int bi = (int)threadIdx.x + q * NT;
            if (TOTAL % NT == 0 || bi < TOTAL) {
                const int line = JFAST ? bi / PER_LINE : bi % NL;
                const int j = JFAST ? bi % PER_LINE : bi / NL;
                float2 v[R];
#pragma unroll
                for (int r = 0; r < R; r++) v[r] = src[fixed_pos<L, NL, JFAST>(line, j + r * PER_LINE)];
                if constexpr (NS > 1) {
#pragma unroll
                    for (int r = 1; r < R; r++) v[r] = fmul_tw<DIR>(v[r], w[q * (R - 1) + r - 1]);
                }
                fdft<R, DIR>(v);
                const int k = j & (NS - 1);
                const int base = (j - k) * R + k;
#pragma unroll
                for (int r = 0; r < R; r++) dst[fixed_pos<L, NL, JFAST>(line, base + r * NS)] = v[r];
            }
        }
        __syncthreads();
        return fixed_stages<DIR, L, NL, NT, MAXR, JFAST, NS * R>(dst, src, w + (NS > 1 ? Q * (R - 1) : 0));
    }
}

// Mesh row of forward/inverse work row a: all rows on the torus; under walls
// rows 0..ph and the last one (cells at -1 wrap there).
__device__ __forceinline__ int mesh_row(const SimDesc& s, int a) {
    if (!s.bounded) return a;
    return a <= s.ph ? a : s.mh - 1;
}

// The forward row transforms of a work item: FROM_MASS, its deposit `mass`
// holds (fixed point, line-major) and converts (S/fft_init.comp) into the
// first buffer; otherwise the first buffer holds the values (a barrier
// passed). Then transform and store the item's lines.
template <int L, bool FROM_MASS, int NT>
__device__ __forceinline__ void rows_forward_fixed(const SimDesc& s, const LineWork& w,
                                                   float2* __restrict__ spec,
                                                   const float2* __restrict__ ftw,
                                                   const uint32_t* mass, float2 (*buf)[ROW_LDS]) {
    constexpr int NL = ROW_LDS / L;
    constexpr int PER = ROW_LDS / NT;
    float2 tws[fixed_twiddle_count<L, NL, NT, 8>() + 1];
    fixed_load_twiddles<L, NL, NT, 8, true>(tws, ftw + fixed_table_offset(L, 8));
    if (FROM_MASS) {
        uint32_t m[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) m[q] = mass[(int)threadIdx.x + q * NT];
#pragma unroll
        for (int q = 0; q < PER; q++)
            buf[0][fswz((int)threadIdx.x + q * NT)] = make_float2((float)m[q] * s.invMassScale, 0.0f);
        __syncthreads();
    }
    const float2* out = fixed_stages<-1, L, NL, NT, 8, true>(buf[0], buf[1], tws);
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int i = (int)threadIdx.x + q * NT;
        const int l = i / L, x = i % L;
        if (l < w.count) spec[s.meshBase + mesh_row(s, w.first + l) * s.mw + x] = out[fswz(i)];
    }
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
__global__ void __launch_bounds__(ROWF_NT) k_fft_rows_forward(const SimDesc* __restrict__ sims,
                                                             const LineWork* __restrict__ work,
                                                             const uint32_t* __restrict__ cellPrefix,
                                                             const float4* __restrict__ cellBody,
                                                             float2* __restrict__ spec,
                                                             const float2* __restrict__ ftw,
                                                             uint32_t* __restrict__ density) {
    // The deposit accumulates in the second transform buffer, which the
    // first stage overwrites only after every lane has converted its values.
    __shared__ float2 buf[2][ROW_LDS];
    uint32_t* mass = reinterpret_cast<uint32_t*>(buf[1]);
    const LineWork w = work[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[w.sim];
    const int L = s.mw;
    for (int i = (int)threadIdx.x; i < ROW_LDS; i += ROWF_NT) mass[i] = 0u;
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
            for (int t = t0 + (int)threadIdx.x; t < t1; t += UNROLL * ROWF_NT) {
                float4 bds[UNROLL];
#pragma unroll
                for (int u = 0; u < UNROLL; u++) bds[u] = cellBody[min(t + u * ROWF_NT, t1 - 1)];
#pragma unroll
                for (int u = 0; u < UNROLL; u++) {
                    if (t + u * ROWF_NT >= t1) continue;
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
    __syncthreads();
    if (DENSITY) {
        for (int i = (int)threadIdx.x; i < w.count * L; i += ROWF_NT) {
            int l = i / L, x = i - l * L;
            density[s.meshBase + mesh_row(s, w.first + l) * s.mw + x] = mass[i];
        }
    }
    switch (L) {
        case 2048: rows_forward_fixed<2048, true, ROWF_NT>(s, w, spec, ftw, mass, buf); break;
        case 1024: rows_forward_fixed<1024, true, ROWF_NT>(s, w, spec, ftw, mass, buf); break;
        case 512: rows_forward_fixed<512, true, ROWF_NT>(s, w, spec, ftw, mass, buf); break;
        case 256: rows_forward_fixed<256, true, ROWF_NT>(s, w, spec, ftw, mass, buf); break;
        case 128: rows_forward_fixed<128, true, ROWF_NT>(s, w, spec, ftw, mass, buf); break;
        case 64: rows_forward_fixed<64, true, ROWF_NT>(s, w, spec, ftw, mass, buf); break;
        case 32: rows_forward_fixed<32, true, ROWF_NT>(s, w, spec, ftw, mass, buf); break;
        default: rows_forward_fixed<16, true, ROWF_NT>(s, w, spec, ftw, mass, buf); break;
    }
}

// The column transforms of a work item of CB <= NL = COL_LDS / H columns of
// length H (interleaved: row y of column c at y * NL + c; a lane past the
// item's columns repeats its last one). The spectrum values of the item's
// nodes load during the forward transform. FORWARD_ONLY: stop after the
// forward transform and store every row (the density spectrum of the
// reference's validateFft, K/GravityPipeline.kt:378-419).
template <bool FORWARD_ONLY, int H>
__device__ __forceinline__ void columns_fixed(const SimDesc& s, const LineWork& w,
                                              float2* __restrict__ spec,
                                              const float2* __restrict__ green,
                                              const float2* __restrict__ ftw,
                                              float2 (*buf)[COL_LDS]) {
    constexpr int NL = COL_LDS / H;
    constexpr int PER = COL_LDS / COL_NT;
    const int CB = w.count, c0 = w.first;
    float2* const mesh = spec + s.meshBase;
    // a dead row (zero padding under walls) loads nothing and stores zero
    float2 v[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int i = (int)threadIdx.x + q * COL_NT;
        const int y = i / NL, c = min(i % NL, CB - 1);
        const bool live = !s.bounded || y <= s.ph || y == H - 1;
        v[q] = live ? mesh[y * s.mw + c0 + c] : make_float2(0.0f, 0.0f);
    }
    float2 tws[fixed_twiddle_count<H, NL, COL_NT, 8>() + 1];
    fixed_load_twiddles<H, NL, COL_NT, 8, false>(tws, ftw + fixed_table_offset(H, 8));
#pragma unroll
    for (int q = 0; q < PER; q++) buf[0][fswz((int)threadIdx.x + q * COL_NT)] = v[q];
    // G(kx, ky) from the table's quarter kx <= mw / 2, ky <= H / 2: a column
    // past mw / 2 takes the conjugate (imaginary part negated), a row past
    // H / 2 the negated conjugate (real part negated).
    float2 gv[PER];
    if (!FORWARD_ONLY) {
        const float2* G = green + s.meshBase;
        const int hx = s.mw >> 1, qw = hx + 1;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int i = (int)threadIdx.x + q * COL_NT;
            const int ky = i / NL, kx = c0 + min(i % NL, CB - 1);
            const int qx = kx > hx ? s.mw - kx : kx, qy = ky > H / 2 ? H - ky : ky;
            gv[q] = G[qy * qw + qx];
        }
    }
    __syncthreads();
    float2* f = fixed_stages<-1, H, NL, COL_NT, 8, false>(buf[0], buf[1], tws);
    if (FORWARD_ONLY) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int i = (int)threadIdx.x + q * COL_NT;
            const int y = i / NL, c = i % NL;
            if (c < CB) mesh[y * s.mw + c0 + c] = f[fswz(i)];
        }
        return;
    }
    // each lane multiplies the values it will not share before the barrier
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int i = (int)threadIdx.x + q * COL_NT;
        const int ky = i / NL, kx = c0 + min(i % NL, CB - 1);
        float2 g = gv[q];
        if (kx > (s.mw >> 1)) g.y = -g.y;
        if (ky > H / 2) g.x = -g.x;
        const int p = fswz(i);
        const float2 z = cmul(f[p], g);
        f[p] = make_float2(z.x * s.invTotalMass, z.y * s.invTotalMass);
    }
    __syncthreads();
    const float2* out = fixed_stages<1, H, NL, COL_NT, 8, false>(f, f == buf[0] ? buf[1] : buf[0], tws);
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int i = (int)threadIdx.x + q * COL_NT;
        const int y = i / NL, c = i % NL;
        const bool live = !s.bounded || y <= s.ph || y == H - 1;
        if (live && c < CB) mesh[y * s.mw + c0 + c] = out[fswz(i)];
    }
}

template <bool FORWARD_ONLY>
// Two workgroups a CU: four waves a SIMD, so at most 128 VGPRs.
__global__ void __launch_bounds__(COL_NT) __attribute__((amdgpu_waves_per_eu(4)))
k_fft_columns(const SimDesc* __restrict__ sims,
                                                        const LineWork* __restrict__ work,
                                                        float2* __restrict__ spec,
                                                        const float2* __restrict__ green,
                                                        const float2* __restrict__ ftw) {
    __shared__ float2 buf[2][COL_LDS];
    const LineWork w = work[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[w.sim];
    switch (s.mh) {
        case 2048: columns_fixed<FORWARD_ONLY, 2048>(s, w, spec, green, ftw, buf); break;
        case 1024: columns_fixed<FORWARD_ONLY, 1024>(s, w, spec, green, ftw, buf); break;
        case 512: columns_fixed<FORWARD_ONLY, 512>(s, w, spec, green, ftw, buf); break;
        case 256: columns_fixed<FORWARD_ONLY, 256>(s, w, spec, green, ftw, buf); break;
        case 128: columns_fixed<FORWARD_ONLY, 128>(s, w, spec, green, ftw, buf); break;
        case 64: columns_fixed<FORWARD_ONLY, 64>(s, w, spec, green, ftw, buf); break;
        case 32: columns_fixed<FORWARD_ONLY, 32>(s, w, spec, green, ftw, buf); break;
        default: columns_fixed<FORWARD_ONLY, 16>(s, w, spec, green, ftw, buf); break;
    }
}

// The inverse row transforms of a work item of lines of length L: a lane's
// ROW_LDS / ROW_NT values and its twiddles load together, the values go to
// LDS, the stages run, the lines the item holds are stored.
template <int L>
__device__ __forceinline__ void rows_inverse_fixed(const SimDesc& s, const LineWork& w,
                                                   float2* __restrict__ spec,
                                                   const float2* __restrict__ ftw,
                                                   float2 (*buf)[ROW_LDS]) {
    constexpr int NL = ROW_LDS / L;
    constexpr int PER = ROW_LDS / ROW_NT;
    float2 tws[fixed_twiddle_count<L, NL, ROW_NT, 8>() + 1];
    float2 v[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int i = (int)threadIdx.x + q * ROW_NT;
        const int l = min(i / L, w.count - 1), x = i % L;
        v[q] = spec[s.meshBase + mesh_row(s, w.first + l) * s.mw + x];
    }
    fixed_load_twiddles<L, NL, ROW_NT, 8, true>(tws, ftw + fixed_table_offset(L, 8));
#pragma unroll
    for (int q = 0; q < PER; q++) buf[0][fswz((int)threadIdx.x + q * ROW_NT)] = v[q];
    __syncthreads();
    const float2* out = fixed_stages<1, L, NL, ROW_NT, 8, true>(buf[0], buf[1], tws);
    // under walls the gather reads columns 0..pw and the last (cells at -1
    // wrap there); the padding columns are not stored
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int i = (int)threadIdx.x + q * ROW_NT;
        const int l = i / L, x = i % L;
        const bool read = !s.bounded || x <= s.pw || x == L - 1;
        if (l < w.count && read) spec[s.meshBase + mesh_row(s, w.first + l) * s.mw + x] = out[fswz(i)];
    }
}

__global__ void __launch_bounds__(ROW_NT) k_fft_rows_inverse(const SimDesc* __restrict__ sims,
                                                             const LineWork* __restrict__ work,
                                                             float2* __restrict__ spec,
                                                             const float2* __restrict__ ftw) {
    __shared__ float2 buf[2][ROW_LDS];
    const LineWork w = work[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[w.sim];
    switch (s.mw) {
        case 2048: rows_inverse_fixed<2048>(s, w, spec, ftw, buf); break;
        case 1024: rows_inverse_fixed<1024>(s, w, spec, ftw, buf); break;
        case 512: rows_inverse_fixed<512>(s, w, spec, ftw, buf); break;
        case 256: rows_inverse_fixed<256>(s, w, spec, ftw, buf); break;
        case 128: rows_inverse_fixed<128>(s, w, spec, ftw, buf); break;
        case 64: rows_inverse_fixed<64>(s, w, spec, ftw, buf); break;
        case 32: rows_inverse_fixed<32>(s, w, spec, ftw, buf); break;
        default: rows_inverse_fixed<16>(s, w, spec, ftw, buf); break;
    }
}

// Adjoint of the force gather (the transposed CIC interpolation): every body
// spreads its float2 value over its four CIC nodes (no mass), with float LDS
// atomics, then the rows are transformed forward as in k_fft_rows_forward.
__global__ void __launch_bounds__(ROW_NT) k_fft_rows_forward_values(
    const SimDesc* __restrict__ sims, const LineWork* __restrict__ work,
    const uint32_t* __restrict__ cellPrefix, const float4* __restrict__ cellBody,
    const float2* __restrict__ value, float2* __restrict__ spec, const float2* __restrict__ ftw) {
    __shared__ float2 buf[2][ROW_LDS];
    const LineWork w = work[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[w.sim];
    const int L = s.mw;
    for (int i = (int)threadIdx.x; i < ROW_LDS; i += ROW_NT) buf[0][i] = make_float2(0.0f, 0.0f);
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
                    float2* node = &buf[0][fswz(l * L + mx)];
                    atomicAdd(&node->x, wgt[k] * v.x);
                    atomicAdd(&node->y, wgt[k] * v.y);
                }
            }
        }
    }
    __syncthreads();
    switch (L) {
        case 2048: rows_forward_fixed<2048, false, ROW_NT>(s, w, spec, ftw, nullptr, buf); break;
        case 1024: rows_forward_fixed<1024, false, ROW_NT>(s, w, spec, ftw, nullptr, buf); break;
        case 512: rows_forward_fixed<512, false, ROW_NT>(s, w, spec, ftw, nullptr, buf); break;
        case 256: rows_forward_fixed<256, false, ROW_NT>(s, w, spec, ftw, nullptr, buf); break;
        case 128: rows_forward_fixed<128, false, ROW_NT>(s, w, spec, ftw, nullptr, buf); break;
        case 64: rows_forward_fixed<64, false, ROW_NT>(s, w, spec, ftw, nullptr, buf); break;
        case 32: rows_forward_fixed<32, false, ROW_NT>(s, w, spec, ftw, nullptr, buf); break;
        default: rows_forward_fixed<16, false, ROW_NT>(s, w, spec, ftw, nullptr, buf); break;
    }
}

void launch_fft_rows_forward_values(hipStream_t st, const SimDesc* sims, const LineWork* work,
                                    int items, const uint32_t* cellPrefix, const float4* cellBody,
                                    const float2* value, float2* spec, const float2* twiddles) {
    if (items == 0) return;
    launch(k_fft_rows_forward_values, dim3(items), dim3(ROW_NT), 0, st, sims, work, cellPrefix,
           cellBody, value, spec, twiddles);
}

void launch_fft_rows_forward(hipStream_t st, const SimDesc* sims, const LineWork* work, int items,
                             const uint32_t* cellPrefix, const float4* cellBody, float2* spec,
                             const float2* twiddles, bool runs, uint32_t* density) {
    if (items == 0) return;
    auto k = density ? (runs ? k_fft_rows_forward<true, true> : k_fft_rows_forward<true, false>)
                     : (runs ? k_fft_rows_forward<false, true> : k_fft_rows_forward<false, false>);
    launch(k, dim3(items), dim3(ROWF_NT), 0, st, sims, work, cellPrefix, cellBody, spec, twiddles,
           density);
}

void launch_fft_columns(hipStream_t st, const SimDesc* sims, const LineWork* work, int items,
                        float2* spec, const float2* green, const float2* twiddles, bool forwardOnly) {
    if (items == 0) return;
    if (forwardOnly)
        launch(k_fft_columns<true>, dim3(items), dim3(COL_NT), 0, st, sims, work, spec, green, twiddles);
    else
        launch(k_fft_columns<false>, dim3(items), dim3(COL_NT), 0, st, sims, work, spec, green, twiddles);
}

void launch_fft_rows_inverse(hipStream_t st, const SimDesc* sims, const LineWork* work, int items,
                             float2* spec, const float2* twiddles) {
    if (items == 0) return;
    launch(k_fft_rows_inverse, dim3(items), dim3(ROW_NT), 0, st, sims, work, spec, twiddles);
}

int fft_twiddle_entries() { return FIXED_TABLE_ENTRIES; }

void fft_twiddle_table(std::vector<float>& out) {
    // every entry is an entry of the FP64-rounded table of 2048: W_L^m is
    // exp(-2 pi i m (2048 / L) / 2048)
    std::vector<float> tw;
    fft_twiddles(TW_LDS, tw);
    out.assign((size_t)2 * FIXED_TABLE_ENTRIES, 0.0f);
    for (int L = 16; L <= 2048; L *= 2) {
        size_t o = (size_t)fixed_table_offset(L, 8);
        for (int ns = 1; ns < L; ns *= fixed_radix_of(L, 8, ns)) {
            const int R = fixed_radix_of(L, 8, ns);
            if (ns == 1) continue;
            for (int r = 1; r < R; r++)
                for (int k = 0; k < ns; k++, o++) {
                    const size_t m = (size_t)r * k * (L / (ns * R)) * (TW_LDS / L);
                    out[2 * o] = tw[2 * m];
                    out[2 * o + 1] = tw[2 * m + 1];
                }
        }
    }
}

int fft_row_lines_per_group(int L) { return L >= ROW_LDS ? 1 : ROW_LDS / L; }
int fft_columns_per_group(int H, int W) {
    int cb = COL_LDS / H;
    return cb < W ? cb : W;
}

}  // namespace gs
