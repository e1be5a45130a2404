// transcript.h -- host-side Fiat-Shamir transcript over Poseidon2 (width 8, rate 4), the
// ChaCha8 stream used for Poseidon2 constants and Ajtai setup, and the CCS matrix digest.
//
// Framing follows neo-transcript/src/poseidon2.rs:20-335: labels and messages are absorbed
// one byte per field element, lengths as single elements, fields via the rate-4 overwrite
// sponge; challenges absorb ONE then permute.  The permutation is Poseidon2 over Goldilocks
// built from ChaCha8Rng::from_seed(b"neo_poseidon2_goldilocks_seed___")
// (neo-ccs/src/crypto/poseidon2_goldilocks.rs:40-43): x^7 S-box, 4+4 external rounds with the
// circ(2,3,1,1) 4x4 block layer, 22 internal rounds.  The internal diagonal is a stand-in drawn
// from the same stream after the round constants (the p3-goldilocks 0.4.1 diagonal constants
// are not part of the reference checkout), so transcript bytes match the CPU restatement
// but not yet the Rust reference.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "field.h"

namespace ns {

struct ChaCha8 {
  uint32_t key[8];
  uint64_t counter = 0;
  uint32_t buf[64];
  int idx = 64;
  explicit ChaCha8(const uint8_t seed[32]);
  static ChaCha8 from_u64(uint64_t state);
  uint32_t next_u32();
  uint64_t next_u64();
  void fill_bytes(uint8_t* dst, size_t len);
  void refill();
};
void chacha8_block(const uint32_t key[8], uint64_t counter, uint32_t out[16]);

struct Poseidon2 {
  int width;
  uint64_t ext_init[4][16];
  uint64_t ext_term[4][16];
  uint64_t internal_rc[22];
  uint64_t diag[16];
  // width-8 AVX-512 path: the 8 external round-constant vectors and diag[0] + 1
  alignas(64) uint64_t vec_rc[8][8];
  uint64_t diag0p1 = 0;
  bool use_avx512 = false;
  void init(int w, ChaCha8& rng);
  void permute(uint64_t* st) const;
  void permute_scalar(uint64_t* st) const;
};
const Poseidon2& transcript_perm();
const Poseidon2& digest_perm();

class Transcript {
 public:
  explicit Transcript(const char* app_label);
  void append_message(const char* label, const uint8_t* msg, size_t len);
  void append_message(const char* label, const char* msg) {
    append_message(label, (const uint8_t*)msg, std::char_traits<char>::length(msg));
  }
  void append_msg_u64(const char* label, uint64_t v);  // msg = v.to_le_bytes()
  void append_fields(const char* label, const uint64_t* fs, size_t n);
  void append_u64s(const char* label, const uint64_t* us, size_t n);
  void append_k(const char* label, K2 x) {
    uint64_t f[2] = {x.c0, x.c1};
    append_fields(label, f, 2);
  }
  uint64_t challenge_field(const char* label);
  void challenge_fields(const char* label, size_t n, uint64_t* out);
  void digest32(uint8_t out[32]);
  K2 challenge_k(const char* l0, const char* l1) {
    K2 r;
    r.c0 = challenge_field(l0);
    r.c1 = challenge_field(l1);
    return r;
  }
  uint64_t permutations() const { return nperm_; }

 private:
  void permute();
  void absorb(uint64_t x);
  void absorb_label(const char* label);
  void absorb_slice(const uint64_t* in, size_t len);
  uint64_t st_[8];
  int absorbed_ = 0;
  uint64_t nperm_ = 0;
  const Poseidon2* perm_;
};

}  // namespace ns
