// The no-allocation form of the space-time operator application: the
// cell-major scratch of the two-pass scatter lives in a fixed device array
// of this library instead of a buffer allocated per operator, so that the
// entry point stfem_stop_vmult_noalloc calls no allocator.
#include "../../include/stfem.h"

#include <mutex>
#include <vector>

#include "capi_common.h"
#include "stfem_host.h"

namespace stfem
{
  namespace
  {
    // Scratch of stfem_stop_vmult_noalloc: one array per GPU, held until the
    // process ends.  Every call on a GPU shares it, so calls that overlap
    // on one GPU race.
    constexpr size_t scratch_bytes = size_t(64) << 20;
    static __device__ __align__(256) unsigned char scratch[scratch_bytes];

    // device address of `scratch` on the current device, looked up once per
    // device (host code must not pass the symbol itself to a kernel)
    void *scratch_address(int device)
    {
      static std::mutex         mtx;
      static std::vector<void *> addr;
      std::lock_guard<std::mutex> lk(mtx);
      if (device >= int(addr.size()))
        addr.resize(device + 1, nullptr);
      if (!addr[device])
        {
          void *p = nullptr;
          STFEM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
          STFEM_REQUIRE(p, "no-alloc scratch: no device address");
          addr[device] = p;
        }
      return addr[device];
    }
  } // namespace

  size_t noalloc_scratch_bytes() { return scratch_bytes; }

  // dst = op src with the cell results in the fixed scratch array
  template <typename T>
  void vmult_noalloc(const SystemMatrix<T> &op, int device, T *dst, const T *src)
  {
    STFEM_HIP(hipSetDevice(device));
    const size_t need = op.cell_buffer_bytes();
    STFEM_REQUIRE(need <= scratch_bytes,
                  "stfem_stop_vmult_noalloc: the operator's cell buffer (" +
                    std::to_string(need) + " bytes) exceeds the fixed scratch of " +
                    std::to_string(scratch_bytes) + " bytes");
    op.vmult_with_cells(dst, src, static_cast<T *>(scratch_address(device)));
  }
  template void vmult_noalloc(const SystemMatrix<float> &, int, float *, const float *);
  template void vmult_noalloc(const SystemMatrix<double> &, int, double *, const double *);
} // namespace stfem
