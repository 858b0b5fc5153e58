// Instantiates the four-wavefront CHR step kernel shapes of the paired-row layout
// (a packed row and its negation share a column entry; chain states in LDS);
// see chr_kernel.hip.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    {4, 1, 1, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 1, false, 1, 0, true>},
    {4, 1, 2, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 2, false, 1, 0, true>},
    {4, 1, 3, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 3, false, 1, 0, true>},
    {4, 1, 4, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 4, false, 1, 0, true>},
    {4, 1, 5, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 5, false, 1, 0, true>},
    {4, 1, 6, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 6, false, 1, 0, true>},
    {4, 1, 7, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 7, false, 1, 0, true>},
    {4, 1, 8, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 8, false, 1, 0, true>},
    {4, 1, 9, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 9, false, 1, 0, true>},
    {4, 1, 10, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 10, false, 1, 0, true>},
    {4, 1, 11, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 11, false, 1, 0, true>},
    {4, 1, 12, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 12, false, 1, 0, true>},
    {4, 1, 14, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 14, false, 1, 0, true>},
    {4, 1, 16, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<4, 1, 16, false, 1, 0, true>},
};
}  // namespace

ChrKernelTable chr_table_w4p() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
