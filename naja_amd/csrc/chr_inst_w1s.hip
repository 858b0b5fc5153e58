// Instantiates the one-wavefront CHR step kernel shapes of the paired-row
// layout on scaled pairs (ChrArgs::scaled; twins of the paired kernels,
// general and coordinate-only, chain states in LDS and in VGPRs); see
// chr_kernel.hip.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    chr_scaled_entry(1, 1, chr_step_kernel<1, 1, 1, false, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 1, true, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 1, false, 2, 2, true, 1, true>,
                     chr_step_kernel<1, 1, 1, true, 2, 2, true, 1, true>),
    chr_scaled_entry(1, 2, chr_step_kernel<1, 1, 2, false, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 2, true, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 2, false, 2, 2, true, 1, true>,
                     chr_step_kernel<1, 1, 2, true, 2, 2, true, 1, true>),
    chr_scaled_entry(1, 3, chr_step_kernel<1, 1, 3, false, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 3, true, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 3, false, 2, 2, true, 1, true>,
                     chr_step_kernel<1, 1, 3, true, 2, 2, true, 1, true>),
    chr_scaled_entry(1, 4, chr_step_kernel<1, 1, 4, false, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 4, true, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 4, false, 2, 2, true, 1, true>,
                     chr_step_kernel<1, 1, 4, true, 2, 2, true, 1, true>),
    chr_scaled_entry(1, 5, chr_step_kernel<1, 1, 5, false, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 5, true, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 5, false, 2, 2, true, 1, true>,
                     chr_step_kernel<1, 1, 5, true, 2, 2, true, 1, true>),
    chr_scaled_entry(1, 6, chr_step_kernel<1, 1, 6, false, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 6, true, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 6, false, 2, 2, true, 1, true>,
                     chr_step_kernel<1, 1, 6, true, 2, 2, true, 1, true>),
    chr_scaled_entry(1, 7, chr_step_kernel<1, 1, 7, false, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 7, true, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 7, false, 2, 2, true, 1, true>,
                     chr_step_kernel<1, 1, 7, true, 2, 2, true, 1, true>),
    chr_scaled_entry(1, 8, chr_step_kernel<1, 1, 8, false, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 8, true, 2, 0, true, 1, true>,
                     chr_step_kernel<1, 1, 8, false, 2, 2, true, 1, true>,
                     chr_step_kernel<1, 1, 8, true, 2, 2, true, 1, true>),
};
}  // namespace

ChrKernelTable chr_table_w1s() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
