// sppm_mtjump.h -- host side of the mt19937 jump-ahead used to split the
// shared camera sample stream (UniformSample, CameraSampler.h:19-27) over
// many workgroups.
//
// A state is a window of the 624 most recent raw (untempered) words
// x[n] .. x[n+623], oldest first.  The next raw word is
//   x[n+624] = x[n+397] ^ mix(x[n], x[n+1])
// and the next output is temper(x[n+624]).  A freshly seeded engine is the
// window of its 624 seed words (the first output twists).  Only the top bit
// of the oldest word takes part in later outputs.
#pragma once
#include <stdint.h>

namespace sppm {

constexpr int MT_WORDS = 624;
constexpr int MT_DEGREE = 19937;  // dimension of the state space
// Raw words a jump reads: x[n + i + j] for i < MT_DEGREE, j < MT_WORDS.
constexpr int MT_JUMP_SEQ = MT_WORDS + MT_DEGREE - 1;

uint32_t mt_mix(uint32_t a, uint32_t b, uint32_t c);
uint32_t mt_temper(uint32_t z);
// Window of std::mt19937(seed) (MT.h:22): the seeding recurrence.
void mt_window_seed(uint32_t seed, uint32_t win[MT_WORDS]);
// Advances a window by `steps` raw words, one at a time.
void mt_window_advance(uint32_t win[MT_WORDS], uint64_t steps);
// Coefficients of t^J mod P(t), P the characteristic polynomial of the
// mt19937 transition: bit i of word i/32 is the coefficient of t^i
// (i < MT_DEGREE; the other bits are zero).
void mt_jump_poly(uint64_t J, uint32_t phi[MT_WORDS]);
// out = the window advanced by J raw words, phi = mt_jump_poly(J):
// out[j] = XOR over i with phi_i = 1 of x[n + i + j].  Equal to the
// sequential result except possibly in the low 31 bits of out[0], which no
// later output reads.
void mt_window_jump(const uint32_t win[MT_WORDS], const uint32_t phi[MT_WORDS], uint32_t out[MT_WORDS]);

}  // namespace sppm
