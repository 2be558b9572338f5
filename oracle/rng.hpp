// ORACLE — test infrastructure only. Nothing in the product links, loads or
// calls this code.
//
// Restatement of the reference random number generators
// (include/dmhm/core/random.hpp:175-514, src/core/random.cpp:17-117):
//  * a 64-bit LCG X <- a X + c mod 2^64 with Knuth's MMIX constants
//    a = 6364136223846793005, c = 1442695040888963407 (random.cpp:17-19).
//    The reference emulates 64-bit products with 16-bit limbs
//    (random.hpp:175-216); native uint64_t arithmetic gives the same residues.
//  * SerialLcg returns the current state and then advances (random.cpp:106-111).
//  * Uniform (0,1]: (upper 32 bits + 1) / 2^32 (random.hpp:238-256).
//  * Box-Muller pairs (random.hpp:280-290) and the column-wise fill of
//    SerialGaussianRandomVectors (random.hpp:418-445), including its index
//    quirk: the last pair of each column is written at rows n-2, n-1 with
//    n = number of COLUMNS (random.hpp:441,443). `ref_quirk` reproduces it.
//  * Initialize seeds both streams with 17 + 21*2^32 (global.cpp:86-93); with
//    one process the "parallel" stream has the serial constants and the same
//    seed but its own state (random.cpp:70-100 with rank 0, commSize 1).
#pragma once
#include <cstdint>
#include <cmath>
#include <vector>

namespace orc {

struct Lcg {
    static constexpr uint64_t A = 6364136223846793005ULL;
    static constexpr uint64_t C = 1442695040888963407ULL;
    uint64_t state;
    // Pre-Initialize value {3,17,19,86} in 16-bit limbs (random.cpp:21).
    Lcg() : state(3ULL | (17ULL << 16) | (19ULL << 32) | (86ULL << 48)) {}
    void seed(uint64_t s) { state = s; }
    uint64_t next() { const uint64_t v = state; state = A * state + C; return v; }
    double uniform() { return ((double)(uint32_t)(next() >> 32) + 1.0) / 4294967296.0; }
    void box_muller(double& x, double& y) {
        const double U = uniform(), V = uniform();
        const double r = std::sqrt(-2.0 * std::log(U));
        x = r * std::cos(2.0 * M_PI * V);
        y = r * std::sin(2.0 * M_PI * V);
    }
    double gaussian_half() {
        const double U = uniform(), V = uniform();
        return std::sqrt(-2.0 * std::log(U)) * std::cos(2.0 * M_PI * V);
    }
    // SerialGaussianRandomVector (random.hpp:364-384); n >= 1.
    void gaussian_vector(int n, double* x) {
        const int numPairs = (n + 1) / 2;
        for (int i = 0; i < numPairs - 1; ++i) box_muller(x[2 * i], x[2 * i + 1]);
        if (n & 1) x[n - 1] = gaussian_half();
        else box_muller(x[n - 2], x[n - 1]);
    }
    // SerialGaussianRandomVectors on an m x n column-major buffer with ld = m
    // (random.hpp:418-445). With ref_quirk the tail of each column goes to rows
    // n-2, n-1 (flat-buffer addressing, writes past the end are dropped).
    void gaussian_vectors(int m, int n, double* A, bool ref_quirk) {
        const int numPairs = (m + 1) / 2;
        const long total = (long)m * n;
        for (int j = 0; j < n; ++j) {
            double* col = A + (long)j * m;
            for (int i = 0; i < numPairs - 1; ++i) box_muller(col[2 * i], col[2 * i + 1]);
            const int last = ref_quirk ? n : m;
            auto at = [&](int row) -> double* {
                const long idx = (long)j * m + row;
                static double sink;
                return (row >= 0 && idx < total) ? &A[idx] : &sink;
            };
            if (m & 1) *at(last - 1) = gaussian_half();
            else { double x, y; box_muller(x, y); *at(last - 2) = x; *at(last - 1) = y; }
        }
    }
};

// The two global streams of the reference after Initialize().
struct Streams {
    Lcg serial, parallel;
    void initialize() {
        const uint64_t s = 17ULL | (21ULL << 32);
        serial.seed(s); parallel.seed(s);
    }
};

} // namespace orc
