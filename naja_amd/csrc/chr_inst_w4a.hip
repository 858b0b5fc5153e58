// Instantiates CHR step kernel shapes (W wavefronts, G chains, R rows per
// lane) for the launch table; see chr_kernel.hip.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    {4, 1, 1, chr_step_kernel<4, 1, 1>},
    {4, 1, 2, chr_step_kernel<4, 1, 2>},
    {4, 1, 3, chr_step_kernel<4, 1, 3>},
    {4, 1, 4, chr_step_kernel<4, 1, 4>},
    {4, 1, 5, chr_step_kernel<4, 1, 5>},
    {4, 1, 6, chr_step_kernel<4, 1, 6>},
    {4, 1, 7, chr_step_kernel<4, 1, 7>},
    {4, 1, 8, chr_step_kernel<4, 1, 8>},
    {4, 1, 9, chr_step_kernel<4, 1, 9>},
    {4, 1, 10, chr_step_kernel<4, 1, 10>},
    {4, 1, 11, chr_step_kernel<4, 1, 11>},
    {4, 1, 12, chr_step_kernel<4, 1, 12>},
};
}  // namespace

ChrKernelTable chr_table_w4a() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
