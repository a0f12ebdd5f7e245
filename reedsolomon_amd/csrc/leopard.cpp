// leopard.cpp — Leopard engines (not yet on the device path).
#include "leopard.hpp"

namespace rsx {

struct LeopardState {};

void LeopardDeleter::operator()(LeopardState *s) const { delete s; }

bool leopard8_shape_ok(int k, int m) {
    int p = 1;
    while (p < m) p <<= 1;
    return k + p <= 256;
}
int leopard_init(rsx_codec *) { return RSX_OK; }
int leopard_encode(rsx_codec *, uint8_t *const *, const size_t *, int) {
    return fail(RSX_ERR_NOT_SUPPORTED, "Leopard engine not built yet");
}
int leopard_verify(rsx_codec *, uint8_t *const *, const size_t *, int, int *) {
    return fail(RSX_ERR_NOT_SUPPORTED, "Leopard engine not built yet");
}
int leopard_reconstruct(rsx_codec *, uint8_t *const *, const size_t *, int, bool, uint8_t *) {
    return fail(RSX_ERR_NOT_SUPPORTED, "Leopard engine not built yet");
}

}  // namespace rsx
