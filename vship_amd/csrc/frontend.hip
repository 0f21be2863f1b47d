// frontend.hip - fused colour front-end + 5-level pyramid + XYB for gfx950.
//
// One workgroup of 256 threads owns a 64x64 level-0 tile of one image; thread
// (tx, ty) owns the 4x4 block at (4tx, 4ty). Per tile, in one pass:
//   raw samples -> range (rangeToFull.hpp:24-55) -> chroma Catmull-Rom
//   upsampling through LDS (chromaUpsample.hpp:28-247) -> matrix + EOTF
//   (YUVToLinRGB.hpp:157-181) -> primaries (primaries.hpp:134-158)
//   -> linear-RGB 2x2 box pyramid with the edge clamp (downsample.hpp:3-19):
//      level 1 and 2 in registers, levels 3 and 4 through LDS
//   -> positive XYB (makeXYB.hpp:6-64) of exactly the planes the score uses.
// A 64x64 level-0 tile maps onto whole 4x4 level-4 blocks, and the edge clamp
// of every level only reads pixels inside the same tile, so tiles are
// independent. The reference instead writes float planes of every level and
// channel back to HBM between ~40 launches per pair.
#include "ssimu2_common.hpp"
#include "color_device.hpp"
#include "kernels.hpp"

namespace vship {

namespace {

struct ChromaStage {
    int ccols, crows;   // staged chroma block incl. 2-sample halo each side
    int cx0, cy0;       // chroma coordinates of staged element (0, 0)
    bool mid;           // horizontally upsampled rows staged too
};

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Segment parameter of full-resolution coordinate X for decimation 2^sub
// (sub >= 1), cosited or centred siting: X maps to segment [k, k+1] at t.
// Cosited: the sample k sits on X = k * 2^sub (kernels *LeftUpscale*,
// *TopUpscale*); centred: on X = k * 2^sub + 2^(sub-1) - 0.5
// (kernels *CenterUpscale*), chromaUpsample.hpp:54-149.
__device__ inline void seg_of(int X, int sub, bool cosited, int& k, float& t, bool& copy)
{
    const int mask = (1 << sub) - 1;
    if (cosited) {
        k = X >> sub;
        const int r = X & mask;
        copy = (r == 0);
        t = (float)r * (1.0f / (float)(1 << sub));
    } else {
        const int off = 1 << (sub - 1);
        k = (X - off) >> sub;  // arithmetic shift: X < off gives k = -1
        const int r = (X - off) & mask;
        copy = false;
        t = ((float)r + 0.5f) * (1.0f / (float)(1 << sub));
    }
}

template <typename StoreT>
__device__ inline void store_row4(StoreT* dst, const float v[4], int xg, int width)
{
#pragma unroll
    for (int i = 0; i < 4; i++)
        if (xg + i < width) dst[i] = store_cast<StoreT>(v[i]);
}

template <int KIND>
__device__ inline float load_kind(const uint8_t* p, uint32_t mask)
{
    if constexpr (KIND == kSampleU8) {
        return (float)p[0];
    } else if constexpr (KIND == kSampleU16) {
        return (float)(((uint32_t)p[0] | ((uint32_t)p[1] << 8)) & mask);
    } else if constexpr (KIND == kSampleHalf) {
        return __half2float(__ushort_as_half((uint16_t)((uint32_t)p[0] | ((uint32_t)p[1] << 8))));
    } else {
        return __uint_as_float((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24));
    }
}

template <typename StoreT, int KIND>
__device__ __attribute__((always_inline)) void front_tile(const FrontJob& job, const FrontColor& fc, const PairGeom& geo, int tile_x, int tile_y,
                           float* lds)
{
    constexpr int kBytes = (KIND == kSampleU8) ? 1 : (KIND == kSampleFloat ? 4 : 2);
    const int t = threadIdx.x;
    const int tx = t & 15, ty = t >> 4;
    const int W = job.width, H = job.height;
    const int x0 = tile_x * kFrontTile + 4 * tx;
    const int y0 = tile_y * kFrontTile + 4 * ty;
    const uint32_t mask = (fc.bits >= 16) ? 0xffffu : ((1u << fc.bits) - 1u);

        float P[3][4][4];  // plane values, then linear R, G, B of the 4x4 block

        // plane 0 (and all planes without subsampling) straight from HBM
        const bool sub = fc.subw || fc.subh;
#pragma unroll
        for (int p = 0; p < 3; p++) {
            if (p > 0 && sub) break;
            const RangeXform rx = make_range(fc.range, fc.bits, p != 0 && fc.yuv);
            const uint8_t* base = job.plane[p];
            const int64_t stride = job.stride[p];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint8_t* row = base + (int64_t)min(y0 + j, H - 1) * stride;
#pragma unroll
                for (int i = 0; i < 4; i++)
                    P[p][j][i] = apply_range(load_kind<KIND>(row + min(x0 + i, W - 1) * kBytes, mask), rx);
            }
        }

        if (sub) {
            // stage both chroma planes (decoded, range-normalised) with a
            // 2-sample halo; clamped coordinates reproduce the reference's
            // edge rules (chromaUpsample.hpp:29-35,41-47)
            ChromaStage cs;
            cs.ccols = (kFrontTile >> fc.subw) + 4;
            cs.crows = (kFrontTile >> fc.subh) + 4;
            cs.cx0 = ((tile_x * kFrontTile) >> fc.subw) - 2;
            cs.cy0 = ((tile_y * kFrontTile) >> fc.subh) - 2;
            cs.mid = fc.subw && fc.subh;
            const int cw = ((W - 1) >> fc.subw) + 1;
            const int ch = ((H - 1) >> fc.subh) + 1;
            const int plane_elems = cs.ccols * cs.crows;
            float* s_chroma = lds;
            float* s_mid = lds + 2 * plane_elems;
            const RangeXform rx = make_range(fc.range, fc.bits, fc.yuv != 0);
            __syncthreads();  // previous item may still read LDS
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const uint8_t* base = job.plane[1 + p];
                const int64_t stride = job.stride[1 + p];
                for (int r = ty; r < cs.crows; r += 16) {
                    const uint8_t* row = base + (int64_t)clampi(cs.cy0 + r, 0, ch - 1) * stride;
                    for (int c = tx; c < cs.ccols; c += 16)
                        s_chroma[p * plane_elems + r * cs.ccols + c] =
                            apply_range(load_kind<KIND>(row + clampi(cs.cx0 + c, 0, cw - 1) * kBytes, mask), rx);
                }
            }
            __syncthreads();
            if (cs.mid) {
                // horizontal pass for every staged chroma row (reference runs
                // it first, into a temporary plane, chromaUpsample.hpp:165-202)
                const int mid_elems = cs.crows * kFrontTile;
                const int X = t & (kFrontTile - 1);
                int k; float tt; bool copy;
                seg_of(tile_x * kFrontTile + X, fc.subw, fc.h_cosited != 0, k, tt, copy);
                const int c = k - cs.cx0;
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    for (int r = t >> 6; r < cs.crows; r += kFrontThreads / kFrontTile) {
                        const float* row = s_chroma + p * plane_elems + r * cs.ccols;
                        s_mid[p * mid_elems + r * kFrontTile + X] =
                            copy ? row[c] : catmull_rom(row[c - 1], row[c], row[c + 1], row[c + 2], tt);
                    }
                }
                __syncthreads();
            }
            // final chroma values of the thread's 4x4 block
#pragma unroll
            for (int p = 0; p < 2; p++) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int Y = y0 + j;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int X = x0 + i;
                        float val;
                        if (fc.subh == 0) {
                            const int r = Y - cs.cy0;
                            const float* row = s_chroma + p * plane_elems + r * cs.ccols;
                            int k; float tt; bool copy;
                            seg_of(X, fc.subw, fc.h_cosited != 0, k, tt, copy);
                            const int c = k - cs.cx0;
                            val = copy ? row[c] : catmull_rom(row[c - 1], row[c], row[c + 1], row[c + 2], tt);
                        } else {
                            int k; float tt; bool copy;
                            seg_of(Y, fc.subh, fc.v_cosited != 0, k, tt, copy);
                            const int r = k - cs.cy0;
                            const float* colbase;
                            int pitch;
                            if (cs.mid) {
                                colbase = s_mid + p * cs.crows * kFrontTile + (X - tile_x * kFrontTile);
                                pitch = kFrontTile;
                            } else {
                                colbase = s_chroma + p * plane_elems + (X - cs.cx0);
                                pitch = cs.ccols;
                            }
                            val = copy ? colbase[r * pitch]
                                       : catmull_rom(colbase[(r - 1) * pitch], colbase[r * pitch],
                                                     colbase[(r + 1) * pitch], colbase[(r + 2) * pitch], tt);
                        }
                        P[1 + p][j][i] = val;
                    }
                }
            }
        }

        // matrix + EOTF + primaries, in place (P becomes linear R, G, B)
#pragma unroll
        for (int j = 0; j < 4; j++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                float R, G, B;
                to_linear_rgb(P[0][j][i], P[1][j][i], P[2][j][i], fc.matrix, fc.transfer, R, G, B);
                if (fc.use_prim) {
                    const float r2 = fmaf(R, fc.prim[0], fmaf(G, fc.prim[1], B * fc.prim[2]));
                    const float g2 = fmaf(R, fc.prim[3], fmaf(G, fc.prim[4], B * fc.prim[5]));
                    const float b2 = fmaf(R, fc.prim[6], fmaf(G, fc.prim[7], B * fc.prim[8]));
                    R = r2; G = g2; B = b2;
                }
                P[0][j][i] = R; P[1][j][i] = G; P[2][j][i] = B;
            }
        }

        StoreT* out = static_cast<StoreT*>(job.out);

        // ---- level 0
        {
            const bool nx = geo.off[0][0] >= 0, ny = geo.off[0][1] >= 0, nb = geo.off[0][2] >= 0;
            if (nx || ny || nb) {
                float vx[4][4], vy[4][4], vb[4][4];
#pragma unroll
                for (int j = 0; j < 4; j++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const XYB q = to_xyb(P[0][j][i], P[1][j][i], P[2][j][i], nb);
                        vx[j][i] = q.x; vy[j][i] = q.y; vb[j][i] = q.b;
                    }
                const int pitch = geo.pitch[0];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (y0 + j >= H) break;
                    const int64_t ro = (int64_t)(y0 + j) * pitch + x0;
                    if (nx) store_row4(out + geo.off[0][0] + ro, vx[j], x0, W);
                    if (ny) store_row4(out + geo.off[0][1] + ro, vy[j], x0, W);
                    if (nb) store_row4(out + geo.off[0][2] + ro, vb[j], x0, W);
                }
            }
        }

        // ---- level 1 (2x2 per thread), edge clamp min(2x+1, w-1)
        float L1[3][2][2];
        {
            const int w0 = W, h0 = H;
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const bool ylast = !(y0 + 2 * j + 1 < h0);
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    const bool xlast = !(x0 + 2 * i + 1 < w0);
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        const float p00 = P[c][2 * j][2 * i];
                        const float p10 = ylast ? p00 : P[c][2 * j + 1][2 * i];
                        const float p01 = xlast ? p00 : P[c][2 * j][2 * i + 1];
                        const float p11 = ylast ? p01 : (xlast ? p10 : P[c][2 * j + 1][2 * i + 1]);
                        float s = p00;
                        s += p10;
                        s += p01;
                        s += p11;
                        L1[c][j][i] = s * 0.25f;
                    }
                }
            }
            const int w1 = geo.w[1], h1 = geo.h[1];
            const int x1 = x0 >> 1, y1 = y0 >> 1;
            const bool nx = geo.off[1][0] >= 0, ny = geo.off[1][1] >= 0, nb = geo.off[1][2] >= 0;
            if (nx || ny || nb) {
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    if (y1 + j >= h1) break;
#pragma unroll
                    for (int i = 0; i < 2; i++) {
                        if (x1 + i >= w1) break;
                        const XYB q = to_xyb(L1[0][j][i], L1[1][j][i], L1[2][j][i], nb);
                        const int64_t o = (int64_t)(y1 + j) * geo.pitch[1] + x1 + i;
                        if (nx) out[geo.off[1][0] + o] = store_cast<StoreT>(q.x);
                        if (ny) out[geo.off[1][1] + o] = store_cast<StoreT>(q.y);
                        if (nb) out[geo.off[1][2] + o] = store_cast<StoreT>(q.b);
                    }
                }
            }
        }

        // ---- level 2 (1 per thread)
        float L2[3];
        {
            const int w1 = geo.w[1], h1 = geo.h[1];
            const int x1 = x0 >> 1, y1 = y0 >> 1;
            const bool xlast = !(x1 + 1 < w1);
            const bool ylast = !(y1 + 1 < h1);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float p00 = L1[c][0][0];
                const float p10 = ylast ? p00 : L1[c][1][0];
                const float p01 = xlast ? p00 : L1[c][0][1];
                const float p11 = ylast ? p01 : (xlast ? p10 : L1[c][1][1]);
                float s = p00;
                s += p10;
                s += p01;
                s += p11;
                L2[c] = s * 0.25f;
            }
            const int x2 = x0 >> 2, y2 = y0 >> 2;
            const bool nx = geo.off[2][0] >= 0, ny = geo.off[2][1] >= 0, nb = geo.off[2][2] >= 0;
            if ((nx || ny || nb) && x2 < geo.w[2] && y2 < geo.h[2]) {
                const XYB q = to_xyb(L2[0], L2[1], L2[2], nb);
                const int64_t o = (int64_t)y2 * geo.pitch[2] + x2;
                if (nx) out[geo.off[2][0] + o] = store_cast<StoreT>(q.x);
                if (ny) out[geo.off[2][1] + o] = store_cast<StoreT>(q.y);
                if (nb) out[geo.off[2][2] + o] = store_cast<StoreT>(q.b);
            }
        }

        // ---- levels 3 and 4 through LDS (16x16 -> 8x8 -> 4x4 per tile)
        float* s_l2 = lds;                 // [3][16][17]
        float* s_l3 = lds + 3 * 16 * 17;   // [3][8][9]
        __syncthreads();                   // chroma staging no longer needed
#pragma unroll
        for (int c = 0; c < 3; c++) s_l2[c * 16 * 17 + ty * 17 + tx] = L2[c];
        __syncthreads();
        if (t < 64) {
            const int lx = t & 7, ly = t >> 3;
            const int x3 = tile_x * 8 + lx, y3 = tile_y * 8 + ly;
            const int x2a = tile_x * 16 + 2 * lx, y2a = tile_y * 16 + 2 * ly;
            const int xb = (x2a + 1 < geo.w[2]) ? 2 * lx + 1 : 2 * lx;
            const int yb = (y2a + 1 < geo.h[2]) ? 2 * ly + 1 : 2 * ly;
            float L3[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float* b2 = s_l2 + c * 16 * 17;
                float s = b2[(2 * ly) * 17 + 2 * lx];
                s += b2[yb * 17 + 2 * lx];
                s += b2[(2 * ly) * 17 + xb];
                s += b2[yb * 17 + xb];
                L3[c] = s * 0.25f;
                s_l3[c * 8 * 9 + ly * 9 + lx] = L3[c];
            }
            const bool nx = geo.off[3][0] >= 0, ny = geo.off[3][1] >= 0, nb = geo.off[3][2] >= 0;
            if ((nx || ny || nb) && x3 < geo.w[3] && y3 < geo.h[3]) {
                const XYB q = to_xyb(L3[0], L3[1], L3[2], nb);
                const int64_t o = (int64_t)y3 * geo.pitch[3] + x3;
                if (nx) out[geo.off[3][0] + o] = store_cast<StoreT>(q.x);
                if (ny) out[geo.off[3][1] + o] = store_cast<StoreT>(q.y);
                if (nb) out[geo.off[3][2] + o] = store_cast<StoreT>(q.b);
            }
        }
        __syncthreads();
        if (t < 16) {
            const int lx = t & 3, ly = t >> 2;
            const int x4 = tile_x * 4 + lx, y4 = tile_y * 4 + ly;
            const int x3a = tile_x * 8 + 2 * lx, y3a = tile_y * 8 + 2 * ly;
            const int xb = (x3a + 1 < geo.w[3]) ? 2 * lx + 1 : 2 * lx;
            const int yb = (y3a + 1 < geo.h[3]) ? 2 * ly + 1 : 2 * ly;
            float L4[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float* b3 = s_l3 + c * 8 * 9;
                float s = b3[(2 * ly) * 9 + 2 * lx];
                s += b3[yb * 9 + 2 * lx];
                s += b3[(2 * ly) * 9 + xb];
                s += b3[yb * 9 + xb];
                L4[c] = s * 0.25f;
            }
            const bool nx = geo.off[4][0] >= 0, ny = geo.off[4][1] >= 0, nb = geo.off[4][2] >= 0;
            if ((nx || ny || nb) && x4 < geo.w[4] && y4 < geo.h[4]) {
                const XYB q = to_xyb(L4[0], L4[1], L4[2], nb);
                const int64_t o = (int64_t)y4 * geo.pitch[4] + x4;
                if (nx) out[geo.off[4][0] + o] = store_cast<StoreT>(q.x);
                if (ny) out[geo.off[4][1] + o] = store_cast<StoreT>(q.y);
                if (nb) out[geo.off[4][2] + o] = store_cast<StoreT>(q.b);
            }
        }
}

template <typename StoreT>
__global__ __launch_bounds__(kFrontThreads)
void frontend_kernel(const FrontJob* __restrict__ jobs, const int* __restrict__ counts,
                     const FrontColor* __restrict__ colors, const PairGeom* __restrict__ geoms)
{
    extern __shared__ float lds[];
    const int total = counts[kCountFrontTiles];
    const int njobs = counts[kCountFrontJobs];
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        // job lookup: largest j with tile_start <= item
        int lo = 0, hi = njobs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobs[mid].tile_start <= item) lo = mid; else hi = mid - 1;
        }
        const FrontJob& job = jobs[lo];
        const FrontColor fc = colors[job.color];
        const PairGeom& geo = geoms[job.pair];
        const int local = item - job.tile_start;
        const int tile_x = local % job.tiles_x;
        const int tile_y = local / job.tiles_x;
        switch (fc.sample) {
        case kSampleU8: front_tile<StoreT, kSampleU8>(job, fc, geo, tile_x, tile_y, lds); break;
        case kSampleU16: front_tile<StoreT, kSampleU16>(job, fc, geo, tile_x, tile_y, lds); break;
        case kSampleHalf: front_tile<StoreT, kSampleHalf>(job, fc, geo, tile_x, tile_y, lds); break;
        default: front_tile<StoreT, kSampleFloat>(job, fc, geo, tile_x, tile_y, lds); break;
        }
        __syncthreads();
    }
}

}  // namespace

size_t frontend_lds_bytes(const FrontColor& fc)
{
    size_t chroma = 0;
    if (fc.subw || fc.subh) {
        const size_t ccols = (kFrontTile >> fc.subw) + 4, crows = (kFrontTile >> fc.subh) + 4;
        chroma = 2 * ccols * crows;
        if (fc.subw && fc.subh) chroma += 2 * crows * kFrontTile;
    }
    const size_t levels = 3 * 16 * 17 + 3 * 8 * 9;
    return sizeof(float) * (chroma > levels ? chroma : levels);
}

size_t frontend_max_lds_bytes()
{
    size_t m = 0;
    for (int subw = 0; subw <= 2; subw++)
        for (int subh = 0; subh <= 1; subh++) {
            FrontColor fc = {};
            fc.subw = subw;
            fc.subh = subh;
            const size_t b = frontend_lds_bytes(fc);
            m = b > m ? b : m;
        }
    return m;
}

template <typename StoreT>
hipError_t launch_frontend(const FrontJob* jobs, const int* counts, int grid, const FrontColor* colors,
                           const PairGeom* geoms, size_t lds_bytes, hipStream_t stream)
{
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(frontend_kernel<StoreT>, dim3(grid), dim3(kFrontThreads), lds_bytes, stream,
                       jobs, counts, colors, geoms);
    return hipGetLastError();
}

template hipError_t launch_frontend<float>(const FrontJob*, const int*, int, const FrontColor*, const PairGeom*, size_t, hipStream_t);
template hipError_t launch_frontend<__half>(const FrontJob*, const int*, int, const FrontColor*, const PairGeom*, size_t, hipStream_t);

}  // namespace vship
