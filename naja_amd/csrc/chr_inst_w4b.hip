// Instantiates CHR step kernel shapes (W wavefronts, G chains, R rows per
// lane) for the launch table; see chr_kernel.hip.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    {4, 1, 14, chr_step_kernel<4, 1, 14>},
    {4, 1, 16, chr_step_kernel<4, 1, 16>},
    {4, 1, 18, chr_step_kernel<4, 1, 18>},
    {4, 1, 20, chr_step_kernel<4, 1, 20>},
    {4, 1, 21, chr_step_kernel<4, 1, 21>},
    {4, 1, 22, chr_step_kernel<4, 1, 22>},
    {4, 1, 24, chr_step_kernel<4, 1, 24>},
    {4, 1, 26, chr_step_kernel<4, 1, 26>},
    {4, 1, 28, chr_step_kernel<4, 1, 28>},
    {4, 1, 30, chr_step_kernel<4, 1, 30>},
    {4, 1, 32, chr_step_kernel<4, 1, 32>},
};
}  // namespace

ChrKernelTable chr_table_w4b() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
