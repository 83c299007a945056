// Frame pre-processing shared by the device kernels (lmz_frame.hip) and the
// host checker: order-preserving keys of pixel values, scipy.ndimage's
// 'reflect' boundary, numpy's median of two middle values, and the two
// window-median algorithms (radix select per output pixel; a sliding
// 256-bin histogram for 8-bit frames).  Everything here is
// __host__ __device__ and free of LDS or wavefront assumptions: the callers
// supply the key fetch and the histogram storage.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "../../include/lmz.h"

namespace lmz {

// scipy.ndimage mode 'reflect' (d c b a | a b c d | d c b a) for any i,
// including windows wider than the image (period 2n).  n < 2^30.
__host__ __device__ inline int reflect_index(int i, int n) {
    if (i >= 0 && i < n) return i;
    const int p = 2 * n;
    int m = i % p;
    if (m < 0) m += p;
    return m < n ? m : p - 1 - m;
}

// Pixel types: storage T, key K (unsigned, same order as the values; every
// NaN maps to the largest key), key bits.
template <int DT> struct Pix;
template <> struct Pix<LMZ_U8> {
    using T = uint8_t;
    using K = uint32_t;
    static constexpr int BITS = 8;
    __host__ __device__ static K key(T v) { return v; }
    __host__ __device__ static T value(K k) { return (T)k; }
    __host__ __device__ static bool isnan(T) { return false; }
};
template <> struct Pix<LMZ_U16> {
    using T = uint16_t;
    using K = uint32_t;
    static constexpr int BITS = 16;
    __host__ __device__ static K key(T v) { return v; }
    __host__ __device__ static T value(K k) { return (T)k; }
    __host__ __device__ static bool isnan(T) { return false; }
};
template <> struct Pix<LMZ_F32> {
    using T = float;
    using K = uint32_t;
    static constexpr int BITS = 32;
    __host__ __device__ static K key(T v) {
        if (v != v) return 0xFFFFFFFFu;
        uint32_t u;
        memcpy(&u, &v, 4);
        return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    }
    __host__ __device__ static T value(K k) {
        const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
        T v;
        memcpy(&v, &u, 4);
        return v;
    }
    __host__ __device__ static bool isnan(T v) { return v != v; }
};
template <> struct Pix<LMZ_F64> {
    using T = double;
    using K = uint64_t;
    static constexpr int BITS = 64;
    __host__ __device__ static K key(T v) {
        if (v != v) return ~0ull;
        uint64_t u;
        memcpy(&u, &v, 8);
        return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
    }
    __host__ __device__ static T value(K k) {
        const uint64_t u = (k & 0x8000000000000000ull)
                               ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
        T v;
        memcpy(&v, &u, 8);
        return v;
    }
    __host__ __device__ static bool isnan(T v) { return v != v; }
};

// numpy.median of an even count: the mean of the two middle values in the
// result type of numpy.mean (float64 for integers, float32 for float32).
template <int DT>
__host__ __device__ inline double mean_of_two(typename Pix<DT>::T a,
                                              typename Pix<DT>::T b) {
    if constexpr (DT == LMZ_F32) {
        const float s = a + b;
        return (double)(s / 2.0f);
    } else {
        return ((double)a + (double)b) / 2.0;
    }
}

// Pixel i of an image of type dt as a double (exact for every type).
__host__ __device__ inline double load_as_double(const void *p, int dt,
                                                 long long i) {
    switch (dt) {
    case LMZ_U8: return (double)((const uint8_t *)p)[i];
    case LMZ_U16: return (double)((const uint16_t *)p)[i];
    case LMZ_F32: return (double)((const float *)p)[i];
    default: return ((const double *)p)[i];
    }
}

// numpy's image / where(background > 0, background, 1.) for one pixel:
// the quotient of the exact operands, rounded once to the output type.
__host__ __device__ inline double normalized(double v, double background) {
    const double safe = background > 0.0 ? background : 1.0;
    return v / safe;
}

__host__ __device__ inline void store_out(void *out, int odt, long long i,
                                          double q) {
    if (odt == LMZ_F32)
        ((float *)out)[i] = (float)q;
    else
        ((double *)out)[i] = q;
}

// Per-thread histogram of 256 bins packed two 16-bit counts per word (bin d
// in word d >> 1, half d & 1).  The storage is supplied by the caller as a
// word accessor: Hist::add(w, v) adds v to word w, Hist::word(w) reads it,
// Hist::clear() zeroes all 128 words.
template <class Hist>
__host__ __device__ inline void hist_add(Hist &h, uint32_t d) {
    h.add(d >> 1, 1u << ((d & 1u) << 4));
}
template <class Hist>
__host__ __device__ inline void hist_sub(Hist &h, uint32_t d) {
    h.add(d >> 1, 0u - (1u << ((d & 1u) << 4)));
}
template <class Hist>
__host__ __device__ inline int hist_get(Hist &h, int d) {
    return (int)((h.word(d >> 1) >> ((d & 1) << 4)) & 0xFFFFu);
}

// First bin b with (count of bins < b) + count[b] > rank; *below gets the
// count of bins < b.
template <class Hist>
__host__ __device__ inline int hist_find(Hist &h, int rank, int *below) {
    int cum = 0;
    for (int w = 0; w < 128; ++w) {
        const uint32_t v = h.word(w);
        const int c0 = (int)(v & 0xFFFFu), c1 = (int)(v >> 16);
        if (cum + c0 > rank) { *below = cum; return 2 * w; }
        cum += c0;
        if (cum + c1 > rank) { *below = cum; return 2 * w + 1; }
        cum += c1;
    }
    *below = cum;
    return 255;   // not reached for rank < total
}

// Count key k in digit (k >> s) & 255 if it agrees with prefix under
// himask; otherwise add 0 (no branch, so that loads can run ahead).
template <class K, class Hist>
__host__ __device__ inline void count_key(Hist &h, K k, K prefix, K himask,
                                          int s) {
    const uint32_t d = (uint32_t)((k >> s) & 0xFFu);
    const uint32_t inc = ((k ^ prefix) & himask) == 0 ? (1u << ((d & 1u) << 4))
                                                      : 0u;
    h.add(d >> 1, inc);
}

// Key of rank `rank` (0-based) among the keys fetch(r, c), r, c in
// [0, size): radix select over 8-bit digits, most significant first.  The
// keys are known to agree above bit top_shift + 8 (they all equal `prefix`
// there), so selection starts at digit top_shift.  Counts fit the 16-bit
// halves for size <= 255.  Keys are fetched eight at a time before they
// are counted.
template <class K, class Fetch, class Hist>
__host__ __device__ inline K window_select(const Fetch &fetch, int size,
                                           int rank, K prefix, int top_shift,
                                           Hist &h) {
    for (int s = top_shift; s >= 0; s -= 8) {
        const K himask = (s + 8 >= (int)(8 * sizeof(K)))
                             ? (K)0 : (K)(~(K)0 << (s + 8));
        h.clear();
        for (int r = 0; r < size; ++r) {
            int c = 0;
            for (; c + 8 <= size; c += 8) {
                K k[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) k[j] = fetch(r, c + j);
#pragma unroll
                for (int j = 0; j < 8; ++j) count_key(h, k[j], prefix, himask, s);
            }
            for (; c < size; ++c) count_key(h, fetch(r, c), prefix, himask, s);
        }
        int below;
        const int b = hist_find(h, rank, &below);
        rank -= below;
        prefix = (prefix & himask) | ((K)b << s);
    }
    return prefix;
}

// Sliding median of 8-bit values down one column of output pixels.  The
// window's histogram is h; `m` is the bin holding the rank-`rank` value and
// `below` the count of values < m.  After adding and removing values (with
// slide_add / slide_remove), slide_settle moves m back onto the rank.
template <class Hist>
__host__ __device__ inline void slide_add(Hist &h, uint32_t v, int m,
                                          int &below) {
    hist_add(h, v);
    below += ((int)v < m);
}
template <class Hist>
__host__ __device__ inline void slide_remove(Hist &h, uint32_t v, int m,
                                             int &below) {
    hist_sub(h, v);
    below -= ((int)v < m);
}
template <class Hist>
__host__ __device__ inline void slide_settle(Hist &h, int rank, int &m,
                                             int &below) {
    while (below > rank) {
        --m;
        below -= hist_get(h, m);
    }
    for (int c = hist_get(h, m); below + c <= rank; c = hist_get(h, m)) {
        below += c;
        ++m;
    }
}

// Two-level sliding median of 16-bit values: a 256-bin histogram H of the
// high bytes of the window, and a 256-bin histogram F of the low bytes of
// the values whose high byte is m, the high byte of the median.  `bc`
// counts values with high byte < m, `bf` values with high byte m and low
// byte < mf, the low byte of the median.  When the median's high byte moves,
// F is rebuilt from the window (refill), which a slowly varying
// background rarely asks for.
struct Slide16 {
    int m = 0, bc = 0, mf = 0, bf = 0;
};

template <class Hist>
__host__ __device__ inline void slide16_add(Hist &H, Hist &F, Slide16 &st,
                                            uint32_t v) {
    const int h = (int)(v >> 8), l = (int)(v & 0xFFu);
    hist_add(H, (uint32_t)h);
    st.bc += (h < st.m);
    if (h == st.m) {
        hist_add(F, (uint32_t)l);
        st.bf += (l < st.mf);
    }
}
template <class Hist>
__host__ __device__ inline void slide16_remove(Hist &H, Hist &F, Slide16 &st,
                                               uint32_t v) {
    const int h = (int)(v >> 8), l = (int)(v & 0xFFu);
    hist_sub(H, (uint32_t)h);
    st.bc -= (h < st.m);
    if (h == st.m) {
        hist_sub(F, (uint32_t)l);
        st.bf -= (l < st.mf);
    }
}
// Median of rank `rank` after updates: settle the high byte; if it moved,
// refill(m) rebuilds F for the new high byte; then settle the low byte.
template <class Hist, class Refill>
__host__ __device__ inline uint32_t slide16_settle(Hist &H, Hist &F,
                                                   Slide16 &st, int rank,
                                                   const Refill &refill) {
    const int m0 = st.m;
    slide_settle(H, rank, st.m, st.bc);
    const int r = rank - st.bc;
    if (st.m != m0) {
        F.clear();
        refill(st.m);
        st.mf = hist_find(F, r, &st.bf);
    } else {
        slide_settle(F, r, st.mf, st.bf);
    }
    return ((uint32_t)st.m << 8) | (uint32_t)st.mf;
}

}  // namespace lmz
