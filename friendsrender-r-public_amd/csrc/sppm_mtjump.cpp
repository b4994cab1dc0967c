// sppm_mtjump.cpp -- jump-ahead for mt19937 over GF(2).
//
// Method (Haramoto, Matsumoto, Nishimura, Panneton, L'Ecuyer, "Efficient
// jump ahead for F2-linear random number generators", INFORMS J. Computing
// 20(3), 2008): the transition T of the generator is linear on the
// 19937-dimensional state space, with characteristic polynomial P.  For a jump
// of J steps, write t^J mod P(t) = sum_i a_i t^i; then T^J s = sum_i a_i T^i s.
// T^i s is the window i words further on, so word j of the jumped window is
// the XOR of x[n + i + j] over the i with a_i = 1.
//
// P is found once per process by Berlekamp-Massey on one bit of the raw word
// sequence.  P is primitive of degree 19937, so any nonzero output bit
// sequence has P as its minimal polynomial.
#include "sppm_mtjump.h"

#include <map>
#include <mutex>
#include <stdexcept>
#include <vector>

namespace sppm {

namespace {
constexpr uint32_t MT_UPPER = 0x80000000u, MT_LOWER = 0x7fffffffu, MT_A = 0x9908b0dfu;

using Words = std::vector<uint64_t>;

inline int bit_of(const Words& v, size_t i) { return (int)((v[i >> 6] >> (i & 63)) & 1u); }
inline void set_bit(Words& v, size_t i) { v[i >> 6] |= 1ull << (i & 63); }
// 64 bits of v starting at bit offset o (v is padded past its last bit)
inline uint64_t get64(const Words& v, size_t o) {
    const size_t w = o >> 6;
    const unsigned b = (unsigned)(o & 63);
    return b ? (v[w] >> b) | (v[w + 1] << (64 - b)) : v[w];
}
// v ^= u << m, over the first nw words of u
inline void xor_shifted(Words& v, const Words& u, size_t nw, size_t m) {
    const size_t ws = m >> 6;
    const unsigned b = (unsigned)(m & 63);
    for (size_t k = 0; k < nw; ++k) {
        const uint64_t x = u[k];
        if (!x) continue;
        v[k + ws] ^= x << b;
        if (b) v[k + ws + 1] ^= x >> (64 - b);
    }
}

// Berlekamp-Massey over GF(2) on bit 31 of x[0], x[1], ...; returns the
// characteristic polynomial (bit d = coefficient of t^d).
Words characteristic_poly() {
    const size_t N = 2 * (size_t)MT_DEGREE + 128;
    // reversed sequence: r_k = s_{N-1-k}
    Words r((N + MT_DEGREE + 256) / 64 + 2, 0);
    {
        std::vector<uint32_t> x(N);
        uint32_t win[MT_WORDS];
        mt_window_seed(5489u, win);
        for (int k = 0; k < MT_WORDS; ++k) x[k] = win[k];
        for (size_t k = MT_WORDS; k < N; ++k) x[k] = mt_mix(x[k - 624], x[k - 623], x[k - 227]);
        for (size_t k = 0; k < N; ++k)
            if (x[k] >> 31) set_bit(r, N - 1 - k);
    }
    const size_t NW = (N + 128) / 64 + 2;
    Words C(NW, 0), B(NW, 0), T;
    C[0] = 1;
    B[0] = 1;
    size_t L = 0, LB = 0, m = 1;
    for (size_t n = 0; n < N; ++n) {
        // discrepancy: sum_{i <= L} C_i s_{n-i}, with s_{n-i} = r_{N-1-n+i}
        const size_t o = N - 1 - n;
        uint64_t d = 0;
        for (size_t k = 0; k <= (L >> 6); ++k) d ^= C[k] & get64(r, o + 64 * k);
        if (!__builtin_parityll(d)) {
            ++m;
            continue;
        }
        if (2 * L <= n) {
            T = C;
            xor_shifted(C, B, (LB >> 6) + 1, m);
            const size_t Lold = L;
            L = n + 1 - L;
            B = T;
            LB = Lold;
            m = 1;
        } else {
            xor_shifted(C, B, (LB >> 6) + 1, m);
            ++m;
        }
    }
    if (L != (size_t)MT_DEGREE) throw std::runtime_error("mt19937 linear complexity is not 19937");
    // P(t) = t^L C(1/t): coefficient of t^(L-i) is C_i
    Words P((MT_DEGREE + 64) / 64 + 1, 0);
    for (size_t i = 0; i <= L; ++i)
        if (bit_of(C, i)) set_bit(P, L - i);
    return P;
}

const Words& char_poly() {
    static std::once_flag once;
    static Words P;
    std::call_once(once, [] { P = characteristic_poly(); });
    return P;
}

constexpr size_t PW = (MT_DEGREE + 64) / 64 + 1;  // words of a reduced polynomial (+ room for one shift)

// v (degree < 2 * MT_DEGREE) mod P, in place
void reduce(Words& v, const std::vector<Words>& Psh, size_t top) {
    for (size_t d = top + 1; d-- > (size_t)MT_DEGREE;) {
        if (!bit_of(v, d)) continue;
        const size_t s = d - MT_DEGREE;
        const Words& p = Psh[s & 63];
        const size_t ws = s >> 6;
        for (size_t k = 0; k < p.size(); ++k) v[k + ws] ^= p[k];
    }
}

uint64_t spread32(uint32_t x) {
    uint64_t v = x;
    v = (v | (v << 16)) & 0x0000ffff0000ffffull;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}

Words jump_poly_words(uint64_t J) {
    const Words& P = char_poly();
    std::vector<Words> Psh(64);
    for (unsigned b = 0; b < 64; ++b) {
        Psh[b].assign(PW + 1, 0);
        xor_shifted(Psh[b], P, P.size(), b);
        while (!Psh[b].empty() && Psh[b].back() == 0) Psh[b].pop_back();
    }
    const size_t SW = 2 * PW + 4;
    Words res(SW, 0), sq(SW, 0);
    res[0] = 1;  // t^0
    int top = 63;
    while (top >= 0 && !((J >> top) & 1)) --top;
    for (int bit = top; bit >= 0; --bit) {
        // res = res^2 mod P (squaring over GF(2) spreads the bits)
        std::fill(sq.begin(), sq.end(), 0);
        for (size_t k = 0; k < PW; ++k) {
            sq[2 * k] = spread32((uint32_t)res[k]);
            sq[2 * k + 1] = spread32((uint32_t)(res[k] >> 32));
        }
        reduce(sq, Psh, 2 * (size_t)MT_DEGREE);
        res.swap(sq);
        if ((J >> bit) & 1) {
            // res *= t
            for (size_t k = PW + 1; k-- > 1;) res[k] = (res[k] << 1) | (res[k - 1] >> 63);
            res[0] <<= 1;
            reduce(res, Psh, MT_DEGREE);
        }
    }
    res.resize(PW);
    return res;
}
}  // namespace

uint32_t mt_mix(uint32_t a, uint32_t b, uint32_t c) {
    const uint32_t y = (a & MT_UPPER) | (b & MT_LOWER);
    return c ^ (y >> 1) ^ ((y & 1u) ? MT_A : 0u);
}

uint32_t mt_temper(uint32_t z) {
    z ^= (z >> 11);
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= (z >> 18);
    return z;
}

void mt_window_seed(uint32_t seed, uint32_t win[MT_WORDS]) {
    win[0] = seed;
    for (int i = 1; i < MT_WORDS; ++i) win[i] = 1812433253u * (win[i - 1] ^ (win[i - 1] >> 30)) + (uint32_t)i;
}

void mt_window_advance(uint32_t win[MT_WORDS], uint64_t steps) {
    // ring of the window; head = oldest word
    uint32_t ring[MT_WORDS];
    for (int i = 0; i < MT_WORDS; ++i) ring[i] = win[i];
    unsigned head = 0;
    for (uint64_t s = 0; s < steps; ++s) {
        const unsigned i1 = head + 1 == MT_WORDS ? 0 : head + 1;
        const unsigned i397 = head + 397 >= MT_WORDS ? head + 397 - MT_WORDS : head + 397;
        ring[head] = mt_mix(ring[head], ring[i1], ring[i397]);
        head = i1;
    }
    for (int i = 0; i < MT_WORDS; ++i) {
        const unsigned k = head + i >= MT_WORDS ? head + i - MT_WORDS : head + i;
        win[i] = ring[k];
    }
}

void mt_jump_poly(uint64_t J, uint32_t phi[MT_WORDS]) {
    static std::mutex mu;
    static std::map<uint64_t, std::vector<uint32_t>> cache;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(J);
        if (it != cache.end()) {
            for (int i = 0; i < MT_WORDS; ++i) phi[i] = it->second[i];
            return;
        }
    }
    const Words w = jump_poly_words(J);
    std::vector<uint32_t> out(MT_WORDS, 0);
    for (int i = 0; i < MT_WORDS; ++i) out[i] = (uint32_t)(w[i >> 1] >> (32 * (i & 1)));
    {
        std::lock_guard<std::mutex> g(mu);
        cache[J] = out;
    }
    for (int i = 0; i < MT_WORDS; ++i) phi[i] = out[i];
}

void mt_window_jump(const uint32_t win[MT_WORDS], const uint32_t phi[MT_WORDS], uint32_t out[MT_WORDS]) {
    std::vector<uint32_t> x(MT_JUMP_SEQ);
    for (int k = 0; k < MT_WORDS; ++k) x[k] = win[k];
    for (int k = MT_WORDS; k < MT_JUMP_SEQ; ++k) x[k] = mt_mix(x[k - 624], x[k - 623], x[k - 227]);
    for (int j = 0; j < MT_WORDS; ++j) out[j] = 0;
    for (int i = 0; i < MT_DEGREE; ++i) {
        if (!((phi[i >> 5] >> (i & 31)) & 1u)) continue;
        for (int j = 0; j < MT_WORDS; ++j) out[j] ^= x[i + j];
    }
}

}  // namespace sppm

// Host utility exported through the C ABI: jump a window by `steps` words.
extern "C" int sppm_mt19937_jump(const uint32_t* state, uint64_t steps, uint32_t* out) {
    if (!state || !out) return -1;
    try {
        uint32_t phi[sppm::MT_WORDS];
        sppm::mt_jump_poly(steps, phi);
        sppm::mt_window_jump(state, phi, out);
    } catch (...) {
        return -1;
    }
    return 0;
}
