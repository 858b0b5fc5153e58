// Instantiates the eight-wavefront CHR step kernel shapes of the paired-row
// layout on scaled pairs (ChrArgs::scaled; twins of the paired kernels,
// chain states in LDS); see chr_kernel.hip.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    chr_scaled_entry(8, 1, chr_step_kernel<8, 1, 1, false, 1, 0, true, 1, true>),
    chr_scaled_entry(8, 2, chr_step_kernel<8, 1, 2, false, 1, 0, true, 1, true>),
    chr_scaled_entry(8, 4, chr_step_kernel<8, 1, 4, false, 1, 0, true, 1, true>),
    chr_scaled_entry(8, 6, chr_step_kernel<8, 1, 6, false, 1, 0, true, 1, true>),
    chr_scaled_entry(8, 8, chr_step_kernel<8, 1, 8, false, 1, 0, true, 1, true>),
    chr_scaled_entry(8, 10, chr_step_kernel<8, 1, 10, false, 1, 0, true, 1, true>),
    chr_scaled_entry(8, 12, chr_step_kernel<8, 1, 12, false, 1, 0, true, 1, true>),
    chr_scaled_entry(8, 14, chr_step_kernel<8, 1, 14, false, 1, 0, true, 1, true>),
    chr_scaled_entry(8, 16, chr_step_kernel<8, 1, 16, false, 1, 0, true, 1, true>),
    chr_scaled_entry(8, 18, chr_step_kernel<8, 1, 18, false, 1, 0, true, 1, true>),
    chr_scaled_entry(8, 20, chr_step_kernel<8, 1, 20, false, 1, 0, true, 1, true>),
    chr_scaled_entry(8, 24, chr_step_kernel<8, 1, 24, false, 1, 0, true, 1, true>),
};
}  // namespace

ChrKernelTable chr_table_w8s() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
