// XLA custom-call targets over the C ABI (host code only), for registration
// with jax.ffi.register_ffi_target(..., platform="ROCM", api_version=0) at the
// reference's call sites (run_hybrid_sphere.py:200-216): the untyped GPU
// custom-call signature
//   void f(hipStream_t stream, void** buffers, const char* opaque, size_t len)
// with the operator's settings in the opaque string as "key=value" fields
// separated by ';' (hexadecimal pointers, decimal integers):
//   ctx=<fsg_context*>;bnd=<fsg_boundary* or 0>;mode=<0|1>;terms=<1|2|3>;nblocks=<n>
// buffers = [U_0 .. U_{n-1}, out_0 .. out_{n-1}] in the solver's block order.
// The signature has no status: a failure is recorded for fsg_last_error()
// and leaves the outputs unwritten.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "../../include/fluid_gfx950.h"

int fsg_fail(int code, const char* fmt, ...);

namespace {

bool field(const char* s, size_t len, const char* key, unsigned long long* out, int base) {
  const size_t kl = strlen(key);
  size_t i = 0;
  while (i < len) {
    size_t j = i;
    while (j < len && s[j] != ';') ++j;
    if (j - i > kl && memcmp(s + i, key, kl) == 0 && s[i + kl] == '=') {
      std::string v(s + i + kl + 1, j - i - kl - 1);
      char* end = nullptr;
      *out = strtoull(v.c_str(), &end, base);
      return end && *end == '\0' && !v.empty();
    }
    i = j + 1;
  }
  return false;
}

}  // namespace

extern "C" {

int fsg_xla_parse(const char* opaque, size_t len, fsg_xla_rhs_config* cfg) {
  if (!opaque || !cfg) return fsg_fail(-10, "null argument");
  unsigned long long ctx = 0, bnd = 0, mode = 0, terms = 0, nb = 0;
  if (!field(opaque, len, "ctx", &ctx, 16) || !field(opaque, len, "bnd", &bnd, 16) ||
      !field(opaque, len, "mode", &mode, 10) || !field(opaque, len, "terms", &terms, 10) ||
      !field(opaque, len, "nblocks", &nb, 10))
    return fsg_fail(-11, "malformed custom-call opaque: need ctx, bnd, mode, terms, nblocks");
  if (!ctx || mode > 1 || terms < 1 || terms > 3 || nb < 1 || nb > 3)
    return fsg_fail(-11, "custom-call opaque out of range");
  cfg->ctx = reinterpret_cast<fsg_context*>(static_cast<uintptr_t>(ctx));
  cfg->boundary = reinterpret_cast<const fsg_boundary*>(static_cast<uintptr_t>(bnd));
  cfg->mode = (int32_t)mode;
  cfg->terms = (int32_t)terms;
  cfg->nblocks = (int32_t)nb;
  return 0;
}

void fsg_xla_rhs(void* stream, void** buffers, const char* opaque, size_t len) {
  fsg_xla_rhs_config c;
  if (!buffers || fsg_xla_parse(opaque, len, &c) != 0) return;
  (void)fsg_residual(c.ctx, const_cast<const void* const*>(buffers), buffers + c.nblocks, c.mode,
                     c.terms, c.boundary, stream);
}

}  // extern "C"
