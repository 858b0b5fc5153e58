// Instantiates the eight-wavefront CHR step kernel shapes of the paired-row layout
// (a packed row and its negation share a column entry; chain states in LDS);
// see chr_kernel.hip. Odd and twice-odd row counts per lane above 16 are
// left out: their per-row buffer offsets no longer fit in VGPRs (R = 24
// holds the same rows without spilling).
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    {8, 1, 1, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<8, 1, 1, false, 1, 0, true>},
    {8, 1, 2, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<8, 1, 2, false, 1, 0, true>},
    {8, 1, 4, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<8, 1, 4, false, 1, 0, true>},
    {8, 1, 6, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<8, 1, 6, false, 1, 0, true>},
    {8, 1, 8, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<8, 1, 8, false, 1, 0, true>},
    {8, 1, 10, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<8, 1, 10, false, 1, 0, true>},
    {8, 1, 12, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<8, 1, 12, false, 1, 0, true>},
    {8, 1, 14, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<8, 1, 14, false, 1, 0, true>},
    {8, 1, 16, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<8, 1, 16, false, 1, 0, true>},
    {8, 1, 18, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<8, 1, 18, false, 1, 0, true>},
    {8, 1, 20, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<8, 1, 20, false, 1, 0, true>},
    {8, 1, 24, nullptr, nullptr, nullptr, nullptr, chr_step_kernel<8, 1, 24, false, 1, 0, true>},
};
}  // namespace

ChrKernelTable chr_table_w8p() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
