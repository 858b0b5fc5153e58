// Instantiates the one-wavefront CHR step kernel shapes of the paired-row
// layout (a packed row and its negation share a column entry), general and
// coordinate-only, with the chain states in LDS and in VGPRs (d <= 128); see
// chr_kernel.hip.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    {1, 1, 1, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<1, 1, 1, false, 2, 0, true>,
     chr_step_kernel<1, 1, 1, true, 2, 0, true>, chr_step_kernel<1, 1, 1, false, 2, 2, true>,
     chr_step_kernel<1, 1, 1, true, 2, 2, true>},
    {1, 1, 2, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<1, 1, 2, false, 2, 0, true>,
     chr_step_kernel<1, 1, 2, true, 2, 0, true>, chr_step_kernel<1, 1, 2, false, 2, 2, true>,
     chr_step_kernel<1, 1, 2, true, 2, 2, true>},
    {1, 1, 3, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<1, 1, 3, false, 2, 0, true>,
     chr_step_kernel<1, 1, 3, true, 2, 0, true>, chr_step_kernel<1, 1, 3, false, 2, 2, true>,
     chr_step_kernel<1, 1, 3, true, 2, 2, true>},
    {1, 1, 4, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<1, 1, 4, false, 2, 0, true>,
     chr_step_kernel<1, 1, 4, true, 2, 0, true>, chr_step_kernel<1, 1, 4, false, 2, 2, true>,
     chr_step_kernel<1, 1, 4, true, 2, 2, true>},
    {1, 1, 5, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<1, 1, 5, false, 2, 0, true>,
     chr_step_kernel<1, 1, 5, true, 2, 0, true>, chr_step_kernel<1, 1, 5, false, 2, 2, true>,
     chr_step_kernel<1, 1, 5, true, 2, 2, true>},
    {1, 1, 6, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<1, 1, 6, false, 2, 0, true>,
     chr_step_kernel<1, 1, 6, true, 2, 0, true>, chr_step_kernel<1, 1, 6, false, 2, 2, true>,
     chr_step_kernel<1, 1, 6, true, 2, 2, true>},
    {1, 1, 7, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<1, 1, 7, false, 2, 0, true>,
     chr_step_kernel<1, 1, 7, true, 2, 0, true>, chr_step_kernel<1, 1, 7, false, 2, 2, true>,
     chr_step_kernel<1, 1, 7, true, 2, 2, true>},
    {1, 1, 8, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<1, 1, 8, false, 2, 0, true>,
     chr_step_kernel<1, 1, 8, true, 2, 0, true>, chr_step_kernel<1, 1, 8, false, 2, 2, true>,
     chr_step_kernel<1, 1, 8, true, 2, 2, true>},
};
}  // namespace

ChrKernelTable chr_table_w1p() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
