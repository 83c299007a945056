// liblmz frame pre-processing for gfx950: background normalization of a
// camera frame resident in device memory (utilities/normalizer.py:11-76),
// the running median of a video's background (utilities/vmedian.py:6-121),
// and the crops that feed a batch straight from the frame
// (analysis/Hologram.py:92-108, analysis/Mask.py:82-105).  C ABI in
// include/lmz.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "lmz_device.h"
#include "lmz_frame.h"

namespace lmz {
namespace {

constexpr int SEL_THREADS = 256;
constexpr int SEL_BLOCKS_PER_CU = 4;

// Device state of a whole-image radix select (the scratch of lmz_median).
struct SelectState {
    uint32_t hist[256];
    unsigned long long prefix;   // key bits resolved so far
    unsigned long long rank;     // rank of the sought key within the prefix set
    unsigned long long below;    // elements below the prefix set
    unsigned long long eq;       // elements equal to the final key
    unsigned long long nan;      // NaN count
    unsigned long long next;     // smallest key above the final key
};

__global__ __launch_bounds__(64) void k_sel_init(SelectState *st,
                                                 unsigned long long k0) {
    for (int i = threadIdx.x; i < 256; i += 64) st->hist[i] = 0u;
    if (threadIdx.x == 0) {
        st->prefix = 0ull;
        st->rank = k0;
        st->below = 0ull;
        st->eq = 0ull;
        st->nan = 0ull;
        st->next = ~0ull;
    }
}

// Histogram of digit (key >> s) & 255 over the keys in the current prefix
// set; the first pass also counts NaNs.  Each thread counts runs of one
// digit before adding to the block's histogram: in the leading digits
// nearly every key of a frame falls in one bin, where atomics would
// serialize.
template <int DT>
__global__ __launch_bounds__(SEL_THREADS) void k_sel_hist(
    const typename Pix<DT>::T *__restrict__ img, long long n,
    SelectState *__restrict__ st, int s, int first) {
    using P = Pix<DT>;
    using K = typename P::K;
    __shared__ uint32_t h[256];
    __shared__ unsigned long long nnan;
    h[threadIdx.x] = 0u;
    if (threadIdx.x == 0) nnan = 0ull;
    __syncthreads();
    const K prefix = (K)st->prefix;
    const K himask = (s + 8 >= (int)(8 * sizeof(K))) ? (K)0
                                                      : (K)(~(K)0 << (s + 8));
    unsigned long long my_nan = 0ull;
    uint32_t run_d = 0u, run_n = 0u;
    for (long long i = (long long)blockIdx.x * SEL_THREADS + threadIdx.x; i < n;
         i += (long long)gridDim.x * SEL_THREADS) {
        const typename P::T v = img[i];
        if (first) my_nan += P::isnan(v);
        const K k = P::key(v);
        if (((k ^ prefix) & himask) == 0) {
            const uint32_t d = (uint32_t)((k >> s) & 0xFFu);
            if (d == run_d) {
                ++run_n;
            } else {
                if (run_n) atomicAdd(&h[run_d], run_n);
                run_d = d;
                run_n = 1u;
            }
        }
    }
    if (run_n) atomicAdd(&h[run_d], run_n);
    if (first && my_nan) atomicAdd(&nnan, my_nan);
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&st->hist[threadIdx.x], h[threadIdx.x]);
    if (first && threadIdx.x == 0 && nnan) atomicAdd(&st->nan, nnan);
}

// Bin of the sought rank; the prefix gains that digit.  Clears the
// histogram for the next pass.
__global__ __launch_bounds__(SEL_THREADS) void k_sel_pick(SelectState *st,
                                                          int s, int last) {
    __shared__ unsigned long long scan[SEL_THREADS];
    const int t = threadIdx.x;
    const unsigned long long c = st->hist[t];
    scan[t] = c;
    __syncthreads();
    for (int off = 1; off < SEL_THREADS; off <<= 1) {
        const unsigned long long v = t >= off ? scan[t - off] : 0ull;
        __syncthreads();
        scan[t] += v;
        __syncthreads();
    }
    const unsigned long long excl = scan[t] - c;
    const unsigned long long rank = st->rank;
    __syncthreads();
    if (c && excl <= rank && rank < excl + c) {
        st->prefix |= (unsigned long long)t << s;
        st->rank = rank - excl;
        st->below += excl;
        if (last) st->eq = c;
    }
    st->hist[t] = 0u;
}

// Smallest key above the selected one, needed when the upper middle
// element of an even count differs from the lower one.
template <int DT>
__global__ __launch_bounds__(SEL_THREADS) void k_sel_next(
    const typename Pix<DT>::T *__restrict__ img, long long n,
    SelectState *__restrict__ st, unsigned long long k1) {
    using P = Pix<DT>;
    if (st->below + st->eq > k1) return;
    const unsigned long long key0 = st->prefix;
    unsigned long long best = ~0ull;
    for (long long i = (long long)blockIdx.x * SEL_THREADS + threadIdx.x; i < n;
         i += (long long)gridDim.x * SEL_THREADS) {
        const unsigned long long k = P::key(img[i]);
        if (k > key0 && k < best) best = k;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(best, off);
        best = o < best ? o : best;
    }
    if ((threadIdx.x & 63) == 0 && best != ~0ull) atomicMin(&st->next, best);
}

template <int DT>
__global__ __launch_bounds__(64) void k_sel_finish(const SelectState *st,
                                                   long long n,
                                                   double *median) {
    using P = Pix<DT>;
    using K = typename P::K;
    if (threadIdx.x != 0) return;
    if (n == 0 || st->nan) {
        *median = __longlong_as_double(0x7FF8000000000000ll);
        return;
    }
    const typename P::T v0 = P::value((K)st->prefix);
    if (n & 1) {
        *median = (double)v0;
        return;
    }
    const unsigned long long k1 = (unsigned long long)(n / 2);
    const typename P::T v1 =
        (st->below + st->eq > k1) ? v0 : P::value((K)st->next);
    *median = mean_of_two<DT>(v0, v1);
}

// image / where(background > 0, background, 1.) with the background a
// device scalar, a host scalar or an image.
__global__ __launch_bounds__(256) void k_divide(
    const void *__restrict__ img, int dt, long long n,
    const double *__restrict__ bg_dev, const void *__restrict__ ref, int rdt,
    double ref_scalar, void *__restrict__ out, int odt) {
    const double bs = bg_dev ? *bg_dev : ref_scalar;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (long long)gridDim.x * 256) {
        const double b = ref ? load_as_double(ref, rdt, i) : bs;
        store_out(out, odt, i, normalized(load_as_double(img, dt, i), b));
    }
}

// A thread's private 256-bin histogram in LDS: word w of lane l at
// base[w * 64 + l], so the 64 lanes of a wavefront touch 64 distinct banks.
struct LdsHist {
    uint32_t *p;   // base + lane
    __device__ void add(uint32_t w, uint32_t v) { atomicAdd(p + w * 64, v); }
    __device__ uint32_t word(int w) const { return p[w * 64]; }
    __device__ void clear() {
        for (int w = 0; w < 128; ++w) p[w * 64] = 0u;
    }
};

template <class T>
__device__ inline void emit(const T *img, int nx, int x, int y, T med,
                            T *bg, void *out, int odt) {
    const long long i = (long long)y * nx + x;
    if (bg) bg[i] = med;
    if (out) store_out(out, odt, i, normalized((double)img[i], (double)med));
}

// Median filter of an 8-bit frame (scipy.ndimage.median_filter, mode
// 'reflect'): one thread per output column walks `rows` rows down, keeping
// the window's 256-bin histogram; a step removes the row leaving the window
// and adds the row entering it (2 * size values instead of size^2) and
// moves the median bin by the change.  The block's 64 + size - 1 columns of
// each row come through LDS (reflected once, in colmap); the next step's
// rows are loaded into registers while the current ones are counted.
constexpr int SEG_MAX = 64 + LMZ_MAX_FILTER_SIZE - 1;
constexpr int SEG_PER_LANE = (SEG_MAX + 63) / 64;

__global__ __launch_bounds__(64) void k_medfilt_u8(
    const uint8_t *__restrict__ img, int ny, int nx, int size, int lo,
    int rows, uint8_t *__restrict__ bg, void *__restrict__ out, int odt) {
    __shared__ uint32_t hist[128 * 64];
    __shared__ uint8_t seg[2][SEG_MAX + 2];
    __shared__ int colmap[SEG_MAX];
    const int lane = threadIdx.x;
    const int x0 = blockIdx.x * 64, x = x0 + lane;
    const int y0 = blockIdx.y * rows;
    if (y0 >= ny) return;                        // uniform over the block
    const int y1 = min(y0 + rows, ny);
    const int w = 63 + size;
    const bool active = x < nx;
    for (int p = lane; p < w; p += 64) colmap[p] = reflect_index(x0 - lo + p, nx);
    __syncthreads();
    uint8_t ra[SEG_PER_LANE], rb[SEG_PER_LANE];
    auto load = [&](uint8_t(&r)[SEG_PER_LANE], int row) {
        const uint8_t *src = img + (long long)reflect_index(row, ny) * nx;
#pragma unroll
        for (int k = 0; k < SEG_PER_LANE; ++k) {
            const int p = lane + 64 * k;
            r[k] = p < w ? src[colmap[p]] : (uint8_t)0;
        }
    };
    auto store = [&](const uint8_t(&r)[SEG_PER_LANE], int buf) {
#pragma unroll
        for (int k = 0; k < SEG_PER_LANE; ++k) {
            const int p = lane + 64 * k;
            if (p < w) seg[buf][p] = r[k];
        }
    };
    LdsHist h{hist + lane};
    if (active) h.clear();
    // the window of row y0, one row at a time
    load(ra, y0 - lo);
    for (int r = 0; r < size; ++r) {
        store(ra, 0);
        __syncthreads();
        if (r + 1 < size) load(ra, y0 - lo + r + 1);
        if (active) {
            const uint8_t *sp = seg[0] + lane;
            int c = 0;
            for (; c + 8 <= size; c += 8) {
                uint32_t v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = sp[c + j];
#pragma unroll
                for (int j = 0; j < 8; ++j) hist_add(h, v[j]);
            }
            for (; c < size; ++c) hist_add(h, sp[c]);
        }
        __syncthreads();
    }
    const int rank = size * size / 2;
    int below = 0, m = 0;
    if (active) {
        m = hist_find(h, rank, &below);
        emit<uint8_t>(img, nx, x, y0, (uint8_t)m, bg, out, odt);
    }
    // step y: row y - 1 - lo leaves, row y - 1 - lo + size enters
    if (y0 + 1 < y1) {
        load(ra, y0 - lo);
        load(rb, y0 - lo + size);
    }
    for (int y = y0 + 1; y < y1; ++y) {
        store(ra, 0);
        store(rb, 1);
        __syncthreads();
        if (y + 1 < y1) {
            load(ra, y - lo);
            load(rb, y - lo + size);
        }
        if (active) {
            const uint8_t *so = seg[0] + lane, *sn = seg[1] + lane;
            int c = 0;
            for (; c + 8 <= size; c += 8) {
                uint32_t vo[8], vn[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    vo[j] = so[c + j];
                    vn[j] = sn[c + j];
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    slide_remove(h, vo[j], m, below);
                    slide_add(h, vn[j], m, below);
                }
            }
            for (; c < size; ++c) {
                slide_remove(h, so[c], m, below);
                slide_add(h, sn[c], m, below);
            }
            slide_settle(h, rank, m, below);
            emit<uint8_t>(img, nx, x, y, (uint8_t)m, bg, out, odt);
        }
        __syncthreads();
    }
}

// Median filter of a 16-bit frame: as k_medfilt_u8, with the two-level
// histogram of Slide16 (high bytes of the window; low bytes of the values
// sharing the median's high byte), refilled from the frame when the
// median's high byte moves.
__global__ __launch_bounds__(64) void k_medfilt_u16(
    const uint16_t *__restrict__ img, int ny, int nx, int size, int lo,
    int rows, uint16_t *__restrict__ bg, void *__restrict__ out, int odt) {
    __shared__ uint32_t histH[128 * 64];
    __shared__ uint32_t histF[128 * 64];
    __shared__ uint16_t seg[2][SEG_MAX + 2];
    __shared__ int colmap[SEG_MAX];
    const int lane = threadIdx.x;
    const int x0 = blockIdx.x * 64, x = x0 + lane;
    const int y0 = blockIdx.y * rows;
    if (y0 >= ny) return;                        // uniform over the block
    const int y1 = min(y0 + rows, ny);
    const int w = 63 + size;
    const bool active = x < nx;
    for (int p = lane; p < w; p += 64) colmap[p] = reflect_index(x0 - lo + p, nx);
    __syncthreads();
    uint16_t ra[SEG_PER_LANE], rb[SEG_PER_LANE];
    auto load = [&](uint16_t(&r)[SEG_PER_LANE], int row) {
        const uint16_t *src = img + (long long)reflect_index(row, ny) * nx;
#pragma unroll
        for (int k = 0; k < SEG_PER_LANE; ++k) {
            const int p = lane + 64 * k;
            r[k] = p < w ? src[colmap[p]] : (uint16_t)0;
        }
    };
    auto store = [&](const uint16_t(&r)[SEG_PER_LANE], int buf) {
#pragma unroll
        for (int k = 0; k < SEG_PER_LANE; ++k) {
            const int p = lane + 64 * k;
            if (p < w) seg[buf][p] = r[k];
        }
    };
    LdsHist H{histH + lane}, F{histF + lane};
    if (active) H.clear();
    load(ra, y0 - lo);
    for (int r = 0; r < size; ++r) {
        store(ra, 0);
        __syncthreads();
        if (r + 1 < size) load(ra, y0 - lo + r + 1);
        if (active) {
            const uint16_t *sp = seg[0] + lane;
            for (int c = 0; c < size; ++c) hist_add(H, (uint32_t)(sp[c] >> 8));
        }
        __syncthreads();
    }
    const int rank = size * size / 2;
    Slide16 st;
    int ycur = y0;
    // low bytes of the window values whose high byte is m, from the frame
    auto refill = [&](int m) {
        for (int r = 0; r < size; ++r) {
            const uint16_t *row = img + (long long)reflect_index(ycur - lo + r, ny) * nx;
            for (int c = 0; c < size; ++c) {
                const uint32_t v = row[colmap[lane + c]];
                if ((int)(v >> 8) == m) hist_add(F, v & 0xFFu);
            }
        }
    };
    if (active) {
        st.m = hist_find(H, rank, &st.bc);
        F.clear();
        refill(st.m);
        st.mf = hist_find(F, rank - st.bc, &st.bf);
        emit<uint16_t>(img, nx, x, y0, (uint16_t)((st.m << 8) | st.mf), bg, out, odt);
    }
    if (y0 + 1 < y1) {
        load(ra, y0 - lo);
        load(rb, y0 - lo + size);
    }
    for (int y = y0 + 1; y < y1; ++y) {
        store(ra, 0);
        store(rb, 1);
        __syncthreads();
        if (y + 1 < y1) {
            load(ra, y - lo);
            load(rb, y - lo + size);
        }
        if (active) {
            const uint16_t *so = seg[0] + lane, *sn = seg[1] + lane;
            for (int c = 0; c < size; ++c) {
                slide16_remove(H, F, st, so[c]);
                slide16_add(H, F, st, sn[c]);
            }
            ycur = y;
            const uint32_t v = slide16_settle(H, F, st, rank, refill);
            emit<uint16_t>(img, nx, x, y, (uint16_t)v, bg, out, odt);
        }
        __syncthreads();
    }
}

// Median filter for 16-bit, float32 and float64 frames: one thread per
// output pixel of an 8 x 8 tile, radix select over the window's keys (8-bit
// digits, starting below the bits all of the tile's keys share).  STAGED:
// the tile's keys with their reflected border sit in LDS (row stride
// `stride`, = 8 mod 32 so the 8 x 8 lanes hit distinct banks); otherwise
// they are read from the frame through reflected row/column tables.
template <int DT, bool STAGED>
__global__ __launch_bounds__(64) void k_medfilt_radix(
    const typename Pix<DT>::T *__restrict__ img, int ny, int nx, int size,
    int lo, int stride, typename Pix<DT>::T *__restrict__ bg,
    void *__restrict__ out, int odt) {
    using P = Pix<DT>;
    using K = typename P::K;
    using T = typename P::T;
    __shared__ uint32_t hist[128 * 64];
    extern __shared__ __align__(16) unsigned char dyn[];
    const int span = 7 + size;
    const int bx = blockIdx.x * 8, by = blockIdx.y * 8;
    const int lane = threadIdx.x;
    // unstaged tables: 64-bit row offsets (frames may exceed 2^31 pixels)
    long long *rowoff = reinterpret_cast<long long *>(dyn);
    int *colidx = reinterpret_cast<int *>(rowoff + span);
    K *tile = reinterpret_cast<K *>(dyn);              // staged keys
    const K kref = P::key(img[(long long)reflect_index(by - lo, ny) * nx +
                              reflect_index(bx - lo, nx)]);
    unsigned long long diff = 0ull;
    if (STAGED) {
        for (int idx = lane; idx < span * span; idx += 64) {
            const int i = idx / span, j = idx - i * span;
            const K k = P::key(img[(long long)reflect_index(by - lo + i, ny) * nx +
                                   reflect_index(bx - lo + j, nx)]);
            tile[i * stride + j] = k;
            diff |= (unsigned long long)(k ^ kref);
        }
    } else {
        for (int i = lane; i < span; i += 64) {
            rowoff[i] = (long long)reflect_index(by - lo + i, ny) * nx;
            colidx[i] = reflect_index(bx - lo + i, nx);
        }
        __syncthreads();
        for (int idx = lane; idx < span * span; idx += 64) {
            const int i = idx / span, j = idx - i * span;
            diff |= (unsigned long long)(P::key(img[rowoff[i] + colidx[j]]) ^
                                         kref);
        }
    }
    for (int off = 32; off > 0; off >>= 1) diff |= __shfl_xor(diff, off);
    __syncthreads();
    const int tx = lane & 7, ty = lane >> 3;
    const int x = bx + tx, y = by + ty;
    if (x >= nx || y >= ny) return;
    K key = kref;
    if (diff) {
        const int hb = 63 - __clzll((long long)diff);
        const int top = (hb / 8) * 8;
        const K prefix = (top + 8 >= (int)(8 * sizeof(K)))
                             ? (K)0 : (K)(kref & (~(K)0 << (top + 8)));
        LdsHist h{hist + lane};
        const int rank = size * size / 2;
        if (STAGED) {
            const K *base = tile + ty * stride + tx;
            auto fetch = [&](int r, int c) { return base[r * stride + c]; };
            key = window_select<K>(fetch, size, rank, prefix, top, h);
        } else {
            auto fetch = [&](int r, int c) {
                return P::key(img[rowoff[ty + r] + colidx[tx + c]]);
            };
            key = window_select<K>(fetch, size, rank, prefix, top, h);
        }
    }
    emit<T>(img, nx, x, y, P::value(key), bg, out, odt);
}

// Median of three 8-bit frames, pixel by pixel (the median buffer of
// utilities/vmedian.py:71-73).
__global__ __launch_bounds__(256) void k_median3_u8(
    const uint8_t *__restrict__ frames, long long n, uint8_t *__restrict__ out) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (long long)gridDim.x * 256) {
        const uint8_t a = frames[i], b = frames[n + i], c = frames[2 * n + i];
        const uint8_t lo = min(a, b), hi = max(a, b);
        out[i] = max(lo, min(hi, c));
    }
}

// Pixels of every feature of a batch read from a frame in device memory:
// dense features copy their box row by row; explicit-coordinate features
// take the box-local pixel index[] of each of their pixels (a Mask's pick)
// and also get their coordinates, relative to the feature's origin, from
// the frame's column and row coordinates xs[], ys[].
__global__ __launch_bounds__(256) void k_gather_frame(
    const lmz_feature *__restrict__ feat, const int32_t *__restrict__ tile_feature,
    const float *__restrict__ frame, int ny, int nx,
    const double *__restrict__ xs, const double *__restrict__ ys,
    const int32_t *__restrict__ box,
    const int32_t *__restrict__ index, long long npix_total,
    float *__restrict__ data, float *__restrict__ coords) {
    const int t = blockIdx.x;
    const int f = tile_feature[t];
    const lmz_feature &F = feat[f];
    const int x0 = box[4 * f], y0 = box[4 * f + 1];
    const int w = box[4 * f + 2], h = box[4 * f + 3];
    const float nanv = __int_as_float(0x7FC00000);
    for (int q = 0; q < LMZ_PPT; ++q) {
        const int j = (t - F.tile_offset) * LMZ_TILE + q * LMZ_THREADS + threadIdx.x;
        if (j >= F.npix) break;
        const long long gi = F.pix_offset + j;
        int li = j;
        if (F.width == 0) li = index ? index[gi] : -1;
        const bool ok = w > 0 && (F.width == 0 || F.width == w) && li >= 0 &&
                        li < w * h;
        const int c = ok ? li % w : 0, r = ok ? li / w : 0;
        const int fx = x0 + c, fy = y0 + r;
        const bool in = ok && fx >= 0 && fx < nx && fy >= 0 && fy < ny;
        data[gi] = in ? frame[(long long)fy * nx + fx] : nanv;
        if (F.width == 0 && coords) {
            coords[gi] = in ? (float)(xs[fx] - F.origin[0]) : nanv;
            coords[npix_total + gi] = in ? (float)(ys[fy] - F.origin[1]) : nanv;
            coords[2 * npix_total + gi] = in ? (float)(0.0 - F.origin[2]) : nanv;
        }
    }
}

int device_cus() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess)
        cus = 256;
    return cus;
}

int grid_for(long long n, int threads) {
    const long long want = (n + threads - 1) / threads;
    return (int)std::max(1LL, std::min(want, (long long)device_cus() * SEL_BLOCKS_PER_CU));
}

bool valid_dtype(int dt) { return dt >= LMZ_U8 && dt <= LMZ_F64; }

template <int DT>
int median_impl(const void *image, long long n, double *median, void *scratch,
                hipStream_t stream) {
    using P = Pix<DT>;
    auto *st = reinterpret_cast<SelectState *>(scratch);
    const auto *img = reinterpret_cast<const typename P::T *>(image);
    const unsigned long long k0 = n > 0 ? (unsigned long long)((n - 1) / 2) : 0ull;
    hipLaunchKernelGGL(k_sel_init, dim3(1), dim3(64), 0, stream, st, k0);
    if (n > 0) {
        const int grid = grid_for(n, SEL_THREADS);
        for (int s = P::BITS - 8; s >= 0; s -= 8) {
            hipLaunchKernelGGL(k_sel_hist<DT>, dim3(grid), dim3(SEL_THREADS), 0,
                               stream, img, n, st, s, s == P::BITS - 8 ? 1 : 0);
            hipLaunchKernelGGL(k_sel_pick, dim3(1), dim3(SEL_THREADS), 0, stream,
                               st, s, s == 0 ? 1 : 0);
        }
        if (!(n & 1))
            hipLaunchKernelGGL(k_sel_next<DT>, dim3(grid), dim3(SEL_THREADS), 0,
                               stream, img, n, st,
                               (unsigned long long)(n / 2));
    }
    hipLaunchKernelGGL(k_sel_finish<DT>, dim3(1), dim3(64), 0, stream, st, n,
                       median);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

int max_lds_per_block() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess)
        return 65536;
    return v;
}

template <int DT>
int medfilt_radix(const void *image, int ny, int nx, int size, void *bgv,
                  void *out, int odt, hipStream_t stream) {
    using P = Pix<DT>;
    using T = typename P::T;
    const int lo = size / 2;
    const int span = 7 + size;
    const int stride = span <= 8 ? 8 : 8 + 32 * ((span - 8 + 31) / 32);
    const size_t hist_bytes = 128 * 64 * sizeof(uint32_t);
    const size_t staged = (size_t)span * stride * sizeof(typename P::K);
    const dim3 grid((nx + 7) / 8, (ny + 7) / 8);
    const auto *img = reinterpret_cast<const T *>(image);
    auto *bg = reinterpret_cast<T *>(bgv);
    bool fits = hist_bytes + staged <= (size_t)max_lds_per_block();
    if (fits && staged > 32768)
        fits = hipFuncSetAttribute(
                   reinterpret_cast<const void *>(&k_medfilt_radix<DT, true>),
                   hipFuncAttributeMaxDynamicSharedMemorySize,
                   (int)staged) == hipSuccess;
    (void)hipGetLastError();
    if (fits) {
        hipLaunchKernelGGL((k_medfilt_radix<DT, true>), grid, dim3(64), staged,
                           stream, img, ny, nx, size, lo, stride, bg, out, odt);
    } else {
        hipLaunchKernelGGL((k_medfilt_radix<DT, false>), grid, dim3(64),
                           span * (sizeof(long long) + sizeof(int)), stream, img,
                           ny, nx, size, lo, stride, bg, out, odt);
    }
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

}  // namespace
}  // namespace lmz

using namespace lmz;

extern "C" {

size_t lmz_median_scratch_bytes(void) { return sizeof(SelectState); }

int lmz_median(const void *image, int32_t dtype, int64_t n, double *median,
               void *scratch, hipStream_t stream) {
    LMZ_REQUIRE(valid_dtype(dtype), "dtype must be LMZ_U8, LMZ_U16, LMZ_F32 or LMZ_F64");
    LMZ_REQUIRE(n >= 0, "n must be >= 0");
    LMZ_REQUIRE(median && scratch, "median and scratch are required");
    LMZ_REQUIRE(image || n == 0, "image is NULL");
    LMZ_REQUIRE(((uintptr_t)scratch & 7) == 0, "scratch must be 8-byte aligned");
    switch (dtype) {
    case LMZ_U8: return median_impl<LMZ_U8>(image, n, median, scratch, stream);
    case LMZ_U16: return median_impl<LMZ_U16>(image, n, median, scratch, stream);
    case LMZ_F32: return median_impl<LMZ_F32>(image, n, median, scratch, stream);
    default: return median_impl<LMZ_F64>(image, n, median, scratch, stream);
    }
}

int lmz_median_filter(const void *image, int32_t dtype, int32_t ny, int32_t nx,
                      int32_t size, void *background, void *out,
                      int32_t out_dtype, hipStream_t stream) {
    LMZ_REQUIRE(valid_dtype(dtype), "dtype must be LMZ_U8, LMZ_U16, LMZ_F32 or LMZ_F64");
    LMZ_REQUIRE(ny > 0 && nx > 0, "ny and nx must be positive");
    LMZ_REQUIRE(size >= 1 && size <= LMZ_MAX_FILTER_SIZE,
                "size must be in [1, LMZ_MAX_FILTER_SIZE]");
    LMZ_REQUIRE(image != nullptr, "image is NULL");
    LMZ_REQUIRE(background || out, "nothing to compute: background and out are NULL");
    LMZ_REQUIRE(!out || out_dtype == LMZ_F32 || out_dtype == LMZ_F64,
                "out_dtype must be LMZ_F32 or LMZ_F64");
    if (dtype == LMZ_U8) {
        // rows per block: 32, halved while the grid has fewer than four
        // workgroups per CU (the initial window costs size^2 per column)
        const int bx = (nx + 63) / 64;
        const int want = 4 * device_cus();
        int rows = 32;
        while (rows > 8 && (long long)bx * ((ny + rows - 1) / rows) < want) rows /= 2;
        const dim3 grid(bx, (ny + rows - 1) / rows);
        hipLaunchKernelGGL(k_medfilt_u8, grid, dim3(64), 0, stream,
                           (const uint8_t *)image, ny, nx, size, size / 2, rows,
                           (uint8_t *)background, out, out_dtype);
        LMZ_HIP(hipGetLastError());
        return LMZ_OK;
    }
    if (dtype == LMZ_U16) {
        const int bx = (nx + 63) / 64;
        const int want = 2 * device_cus();
        int rows = 32;
        while (rows > 8 && (long long)bx * ((ny + rows - 1) / rows) < want) rows /= 2;
        const dim3 grid(bx, (ny + rows - 1) / rows);
        hipLaunchKernelGGL(k_medfilt_u16, grid, dim3(64), 0, stream,
                           (const uint16_t *)image, ny, nx, size, size / 2, rows,
                           (uint16_t *)background, out, out_dtype);
        LMZ_HIP(hipGetLastError());
        return LMZ_OK;
    }
    switch (dtype) {
    case LMZ_F32:
        return medfilt_radix<LMZ_F32>(image, ny, nx, size, background, out,
                                      out_dtype, stream);
    default:
        return medfilt_radix<LMZ_F64>(image, ny, nx, size, background, out,
                                      out_dtype, stream);
    }
}

int lmz_normalize(const void *image, int32_t dtype, int64_t n,
                  const double *background_scalar, const void *reference,
                  int32_t ref_dtype, double ref_scalar, void *out,
                  int32_t out_dtype, hipStream_t stream) {
    LMZ_REQUIRE(valid_dtype(dtype), "dtype must be LMZ_U8, LMZ_U16, LMZ_F32 or LMZ_F64");
    LMZ_REQUIRE(n >= 0, "n must be >= 0");
    LMZ_REQUIRE((image && out) || n == 0, "image and out are required");
    LMZ_REQUIRE(out_dtype == LMZ_F32 || out_dtype == LMZ_F64,
                "out_dtype must be LMZ_F32 or LMZ_F64");
    LMZ_REQUIRE(!(background_scalar && reference),
                "give one of background_scalar and reference");
    LMZ_REQUIRE(!reference || valid_dtype(ref_dtype),
                "ref_dtype must be LMZ_U8, LMZ_U16, LMZ_F32 or LMZ_F64");
    if (n == 0) return LMZ_OK;
    hipLaunchKernelGGL(k_divide, dim3(grid_for(n, 256)), dim3(256), 0, stream,
                       image, dtype, (long long)n, background_scalar, reference,
                       ref_dtype, ref_scalar, out, out_dtype);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

int lmz_median3_u8(const uint8_t *frames, int64_t n, uint8_t *out,
                   hipStream_t stream) {
    LMZ_REQUIRE(n >= 0, "n must be >= 0");
    LMZ_REQUIRE((frames && out) || n == 0, "frames and out are required");
    if (n == 0) return LMZ_OK;
    hipLaunchKernelGGL(k_median3_u8, dim3(grid_for(n, 256)), dim3(256), 0,
                       stream, frames, (long long)n, out);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

int lmz_gather_frame(const lmz_batch *b, const float *frame, int32_t ny,
                     int32_t nx, const double *xs, const double *ys,
                     const int32_t *box, const int32_t *index, float *data,
                     float *coords, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(frame && box && data, "frame, box and data are required");
    LMZ_REQUIRE(ny > 0 && nx > 0, "ny and nx must be positive");
    LMZ_REQUIRE(b->all_dense || (index && coords && xs && ys),
                "explicit-coordinate features need index, coords, xs and ys");
    hipLaunchKernelGGL(k_gather_frame, dim3(b->ntiles), dim3(LMZ_THREADS), 0,
                       stream, b->d_feat, b->d_tile_feature, frame, ny, nx, xs,
                       ys, box, index, (long long)b->npix_total, data, coords);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

}  // extern "C"
