// ck_internal.hpp -- declarations shared by the host and device halves of
// libcuckoo_mc.so.  Not part of the public ABI (include/cuckoo_mc.h is).
#pragma once
#include "cuckoo_mc.h"

#include <cstdint>
#include <string>

namespace ck {

const uint8_t* sbox();    // FIPS-197 S-box, computed from GF(2^8)
const uint32_t* te0();    // T0[x] = {2S(x), S(x), S(x), 3S(x)}, MSB first

void expand_words(const uint32_t w4[4], uint32_t rk[44]);
void expand_key(uint64_t lo, uint64_t hi, uint32_t rk[44]);
void host_aes(const uint32_t rk[44], const uint32_t in[4], uint32_t out[4]);
void batch_key(uint64_t seed, uint64_t id, uint32_t rk[44]);
float original_log2(uint64_t n);
int derive_params(uint32_t mode, uint32_t pnum, uint32_t ell_i, uint32_t k, uint32_t buckets, uint32_t kick,
                  uint32_t stash, uint32_t input_mode, ck_params* out);

int fail(int code, const std::string& msg);
// CK_OK when *p is exactly what ck_params_derive returns for its inputs.
int check_params(const ck_params* p);
void set_error(const std::string& e);
const char* last_error();

}  // namespace ck
