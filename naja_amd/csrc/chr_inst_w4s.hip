// Instantiates the four-wavefront CHR step kernel shapes of the paired-row
// layout on scaled pairs (ChrArgs::scaled; twins of the paired kernels,
// chain states in LDS); see chr_kernel.hip.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    chr_scaled_entry(4, 1, chr_step_kernel<4, 1, 1, false, 1, 0, true, 1, true>),
    chr_scaled_entry(4, 2, chr_step_kernel<4, 1, 2, false, 1, 0, true, 1, true>),
    chr_scaled_entry(4, 3, chr_step_kernel<4, 1, 3, false, 1, 0, true, 1, true>),
    chr_scaled_entry(4, 4, chr_step_kernel<4, 1, 4, false, 1, 0, true, 1, true>),
    chr_scaled_entry(4, 5, chr_step_kernel<4, 1, 5, false, 1, 0, true, 1, true>),
    chr_scaled_entry(4, 6, chr_step_kernel<4, 1, 6, false, 1, 0, true, 1, true>),
    chr_scaled_entry(4, 7, chr_step_kernel<4, 1, 7, false, 1, 0, true, 1, true>),
    chr_scaled_entry(4, 8, chr_step_kernel<4, 1, 8, false, 1, 0, true, 1, true>),
    chr_scaled_entry(4, 9, chr_step_kernel<4, 1, 9, false, 1, 0, true, 1, true>),
    chr_scaled_entry(4, 10, chr_step_kernel<4, 1, 10, false, 1, 0, true, 1, true>),
    chr_scaled_entry(4, 11, chr_step_kernel<4, 1, 11, false, 1, 0, true, 1, true>),
    chr_scaled_entry(4, 12, chr_step_kernel<4, 1, 12, false, 1, 0, true, 1, true>),
    chr_scaled_entry(4, 14, chr_step_kernel<4, 1, 14, false, 1, 0, true, 1, true>),
    chr_scaled_entry(4, 16, chr_step_kernel<4, 1, 16, false, 1, 0, true, 1, true>),
};
}  // namespace

ChrKernelTable chr_table_w4s() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
