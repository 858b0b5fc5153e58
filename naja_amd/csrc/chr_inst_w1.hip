// Instantiates CHR step kernel shapes (W wavefronts, G chains, R rows per
// lane) for the launch table; see chr_kernel.hip. One-wavefront kernels
// prefetch their columns two steps ahead, with a coordinate-only twin, each
// also with the chain states in VGPRs for d <= 128.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    {1, 1, 1, chr_step_kernel<1, 1, 1, false, 2>, chr_step_kernel<1, 1, 1, true, 2>,
     chr_step_kernel<1, 1, 1, false, 2, 2>, chr_step_kernel<1, 1, 1, true, 2, 2>},
    {1, 1, 2, chr_step_kernel<1, 1, 2, false, 2>, chr_step_kernel<1, 1, 2, true, 2>,
     chr_step_kernel<1, 1, 2, false, 2, 2>, chr_step_kernel<1, 1, 2, true, 2, 2>},
    {1, 1, 3, chr_step_kernel<1, 1, 3, false, 2>, chr_step_kernel<1, 1, 3, true, 2>,
     chr_step_kernel<1, 1, 3, false, 2, 2>, chr_step_kernel<1, 1, 3, true, 2, 2>},
    {1, 1, 4, chr_step_kernel<1, 1, 4, false, 2>, chr_step_kernel<1, 1, 4, true, 2>,
     chr_step_kernel<1, 1, 4, false, 2, 2>, chr_step_kernel<1, 1, 4, true, 2, 2>},
    {1, 1, 5, chr_step_kernel<1, 1, 5, false, 2>, chr_step_kernel<1, 1, 5, true, 2>,
     chr_step_kernel<1, 1, 5, false, 2, 2>, chr_step_kernel<1, 1, 5, true, 2, 2>},
    {1, 1, 6, chr_step_kernel<1, 1, 6, false, 2>, chr_step_kernel<1, 1, 6, true, 2>,
     chr_step_kernel<1, 1, 6, false, 2, 2>, chr_step_kernel<1, 1, 6, true, 2, 2>},
    {1, 1, 7, chr_step_kernel<1, 1, 7, false, 2>, chr_step_kernel<1, 1, 7, true, 2>,
     chr_step_kernel<1, 1, 7, false, 2, 2>, chr_step_kernel<1, 1, 7, true, 2, 2>},
    {1, 1, 8, chr_step_kernel<1, 1, 8, false, 2>, chr_step_kernel<1, 1, 8, true, 2>,
     chr_step_kernel<1, 1, 8, false, 2, 2>, chr_step_kernel<1, 1, 8, true, 2, 2>},
};
}  // namespace

ChrKernelTable chr_table_w1() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
