// One (precision, adjoint, normals source) family of path kernels,
// instantiated for every model, product and LDS layout.  Built once per
// family with -DCF_INST_REAL=0|1 (FP64, FP32), -DCF_INST_ADJ=0|1|2
// (cfk::AdjointKind) and -DCF_INST_SRC=0|1|2|3|4 (Sobol, mrg32k3a, Sobol
// with the itemized-risk modes: checkpoint family only, Sobol kernels of
// packed launches, mrg32k3a with the itemized-risk modes), so the families
// compile in parallel.
#include <type_traits>

#if !defined(CF_INST_REAL) || !defined(CF_INST_ADJ) || !defined(CF_INST_SRC)
#error "CF_INST_REAL, CF_INST_ADJ and CF_INST_SRC must be defined"
#endif
#if (CF_INST_SRC == 2 || CF_INST_SRC == 4) && CF_INST_ADJ != 1
#error "the itemized-risk kernels are checkpoint kernels"
#endif

// FP32 pricing kernels run two 1024-thread blocks per CU (LDS allows it for
// the benchmark grids).  Measured on gfx950: at 92 scalar registers only one
// block stayed resident (half the waves); at most 80 keeps both.
#if CF_INST_REAL == 1 && CF_INST_ADJ == 0
#define CF_PATH_KERNEL_ATTR __attribute__((amdgpu_num_sgpr(80)))
#endif

#include "cf_path_kernel.h"

namespace cfk {
namespace {

using Real = std::conditional_t<CF_INST_REAL == 0, double, float>;
constexpr int kAdj = CF_INST_ADJ;
// CF_INST_SRC: 0 Sobol, 1 mrg32k3a, 2 Sobol with the itemized-risk
// modes (forward only keeping end states, reverse only from them), 3 Sobol
// kernels of packed launches (path_kernel_packed: the same body; with
// multiply-adds fused only within an expression it rounds as the single
// kernel does), 4 mrg32k3a with the itemized-risk modes
constexpr bool kMrg = CF_INST_SRC == 1 || CF_INST_SRC == 4;
constexpr bool kItem = CF_INST_SRC == 2 || CF_INST_SRC == 4;
constexpr bool kPacked = CF_INST_SRC == 3;

template <int MODEL, int PRD>
const void* ptr(bool lean)
{
    if constexpr (kPacked)
        return lean ? reinterpret_cast<const void*>(&path_kernel_packed<Real, MODEL, PRD, kAdj, true>)
                    : reinterpret_cast<const void*>(&path_kernel_packed<Real, MODEL, PRD, kAdj, false>);
    else
        return lean ? reinterpret_cast<const void*>(&path_kernel<Real, MODEL, PRD, kAdj, true, kMrg, kItem>)
                    : reinterpret_cast<const void*>(&path_kernel<Real, MODEL, PRD, kAdj, false, kMrg, kItem>);
}

template <int MODEL>
const void* by_product(int prd, bool lean)
{
    switch (prd) {
    case PRD_EUROPEAN: return ptr<MODEL, PRD_EUROPEAN>(lean);
    case PRD_BARRIER: return ptr<MODEL, PRD_BARRIER>(lean);
    case PRD_EUROPEANS: return ptr<MODEL, PRD_EUROPEANS>(lean);
    case PRD_BOOK: return ptr<MODEL, PRD_BOOK>(lean);
    }
    return nullptr;
}

}  // namespace

const void* CF_PICK_NAME(CF_INST_REAL, CF_INST_ADJ, CF_INST_SRC)(int model, int prd, bool lean)
{
    return model == MODEL_DUPIRE ? by_product<MODEL_DUPIRE>(prd, lean) : by_product<MODEL_BS>(prd, lean);
}

}  // namespace cfk
