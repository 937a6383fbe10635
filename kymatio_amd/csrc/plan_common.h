// Host-side helpers shared by the 1D and 2D plans: error state, FFT length
// tables (twiddles in double, digit-reversal maps) and a relocatable blob
// of plan constants that is copied once to the device.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstring>
#include <string>
#include <vector>

#include "fft_core.h"

enum KymStatus {
    KYM_OK = 0,
    KYM_ERR_TYPE = 1,     // maps to TypeError
    KYM_ERR_RUNTIME = 2,  // maps to RuntimeError
    KYM_ERR_VALUE = 3,    // maps to ValueError
    KYM_ERR_HIP = 4,      // HIP runtime failure
    KYM_ERR_UNSUPPORTED = 5,
};

void kym_set_error(const std::string& msg);

struct KymError {
    int code;
    std::string msg;
};

// Digit-reversal map for DIT radices rad[0..nrad): rev[p] = natural index
// held at storage position p (the permutation a DIT pass sequence consumes
// and the reverse-order DIF sequence produces).
inline void kym_rev_map(int n, const int* rad, int nrad, int* rev) {
    // recursive structure: rev_K(p) = p / (n/R_K) + R_K * rev_{K-1}(p mod (n/R_K))
    std::vector<int> cur(1, 0);
    int m = 1;
    for (int q = 0; q < nrad; ++q) {
        const int R = rad[q] > 1 ? rad[q] : 1;
        std::vector<int> nxt((size_t)m * R);
        for (int p = 0; p < m * R; ++p) nxt[p] = p / m + R * cur[p % m];
        cur.swap(nxt);
        m *= R;
    }
    for (int p = 0; p < n; ++p) rev[p] = cur[p];
}

// Twiddles exp(-2 pi i t / n), evaluated in long double and rounded once.
inline void kym_twiddles(int n, double* re, double* im) {
    for (int t = 0; t < n; ++t) {
        long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
        re[t] = (double)std::cos(ang);
        im[t] = -(double)std::sin(ang);
    }
    // exact values at the quarter points
    if (n % 4 == 0) {
        re[n / 4] = 0.0; im[n / 4] = -1.0;
        re[3 * n / 4] = 0.0; im[3 * n / 4] = 1.0;
    }
    if (n % 2 == 0) { re[n / 2] = -1.0; im[n / 2] = 0.0; }
    re[0] = 1.0; im[0] = 0.0;
}

// Growable byte blob with 256-byte aligned sections.  Pointer fields inside
// it are written as offsets from the blob start and listed in `reloc`;
// relocate(base) turns them into addresses once the blob's final home
// (device or host memory) is known.
struct Blob {
    std::vector<char> bytes;
    std::vector<size_t> reloc;
    size_t add(size_t nbytes) {
        size_t off = (bytes.size() + 255) & ~(size_t)255;
        bytes.resize(off + nbytes);
        return off;
    }
    template <class X> X* at(size_t off) { return reinterpret_cast<X*>(bytes.data() + off); }
    void relocate(uintptr_t base) {
        for (size_t off : reloc) {
            uintptr_t v;
            std::memcpy(&v, bytes.data() + off, sizeof(v));
            v += base;
            std::memcpy(bytes.data() + off, &v, sizeof(v));
        }
        reloc.clear();
    }
};

// Lengths used by a plan: one FftLen per distinct n with twiddle/rev tables.
template <class T>
struct LenTable {
    std::vector<int> ns;
    int max_pow2 = 16;  // largest power-of-two radix (tuning)
    int index_of(int n) const {
        for (size_t i = 0; i < ns.size(); ++i)
            if (ns[i] == n) return (int)i;
        return -1;
    }
    int add(int n) {
        int i = index_of(n);
        if (i >= 0) return i;
        ns.push_back(n);
        return (int)ns.size() - 1;
    }
    // Writes FftLen array + tables into the blob; returns offset of the array.
    bool emit(Blob& b, size_t* lens_off, std::string* err) const {
        std::vector<FftLen<T>> fl(ns.size());
        std::vector<size_t> tw_off(ns.size()), rev_off(ns.size()), irev_off(ns.size());
        for (size_t i = 0; i < ns.size(); ++i) {
            const int n = ns[i];
            FftLen<T>& f = fl[i];
            f.n = n;
            if (!kym_factor(n, f.rad, &f.nrad, max_pow2)) {
                *err = "FFT length " + std::to_string(n) + " has a prime factor above 31 (unsupported)";
                return false;
            }
            tw_off[i] = b.add(sizeof(cpx<T>) * n);
            rev_off[i] = b.add(sizeof(int) * n);
            irev_off[i] = b.add(sizeof(int) * n);
            std::vector<double> re(n), im(n);
            kym_twiddles(n, re.data(), im.data());
            cpx<T>* tw = b.at<cpx<T>>(tw_off[i]);
            for (int t = 0; t < n; ++t) tw[t] = {(T)re[t], (T)im[t]};
            kym_rev_map(n, f.rad, f.nrad, b.at<int>(rev_off[i]));
            const int* rv = b.at<int>(rev_off[i]);
            int* iv = b.at<int>(irev_off[i]);
            for (int t = 0; t < n; ++t) iv[rv[t]] = t;
        }
        *lens_off = b.add(sizeof(FftLen<T>) * ns.size());
        FftLen<T>* dst = b.at<FftLen<T>>(*lens_off);
        for (size_t i = 0; i < ns.size(); ++i) {
            fl[i].tw = reinterpret_cast<const cpx<T>*>(tw_off[i]);
            fl[i].rev = reinterpret_cast<const int*>(rev_off[i]);
            fl[i].irev = reinterpret_cast<const int*>(irev_off[i]);
            dst[i] = fl[i];
            const size_t rec = *lens_off + i * sizeof(FftLen<T>);
            b.reloc.push_back(rec + offsetof(FftLen<T>, tw));
            b.reloc.push_back(rec + offsetof(FftLen<T>, rev));
            b.reloc.push_back(rec + offsetof(FftLen<T>, irev));
        }
        return true;
    }
};
