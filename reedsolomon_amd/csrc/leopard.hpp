// leopard.hpp — the Leopard GF(2^16) / GF(2^8) engines (leopard.go, leopard8.go).
#pragma once

#include <cstddef>
#include <cstdint>

#include "codec.hpp"

namespace rsx {

// k + ceilPow2(m) <= 256: the shapes the reference's GF(2^8) tables cover.
bool leopard8_shape_ok(int k, int m);
// Shapes inside the tables of the reference (see leopard.cpp).
bool leopard_shape_ok(int bits, int k, int m);
int leopard_init(rsx_codec *c);
int leopard_encode(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards);
int leopard_verify(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards, int *ok);
int leopard_reconstruct(rsx_codec *c, uint8_t *const *shards, const size_t *lens, int nshards,
                        bool recover_all, uint8_t *filled);

}  // namespace rsx
