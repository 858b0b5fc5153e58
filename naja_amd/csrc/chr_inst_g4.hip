// Instantiates CHR step kernel shapes (W wavefronts, G chains, R rows per
// lane) for the launch table; see chr_kernel.hip.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    {1, 4, 1, chr_step_kernel<1, 4, 1>},
    {1, 4, 2, chr_step_kernel<1, 4, 2>},
    {1, 4, 3, chr_step_kernel<1, 4, 3>},
    {1, 4, 4, chr_step_kernel<1, 4, 4>},
    {1, 4, 5, chr_step_kernel<1, 4, 5>},
    {1, 4, 6, chr_step_kernel<1, 4, 6>},
    {1, 4, 7, chr_step_kernel<1, 4, 7>},
    {1, 4, 8, chr_step_kernel<1, 4, 8>},
    {4, 4, 1, chr_step_kernel<4, 4, 1>},
    {4, 4, 2, chr_step_kernel<4, 4, 2>},
    {4, 4, 3, chr_step_kernel<4, 4, 3>},
    {4, 4, 4, chr_step_kernel<4, 4, 4>},
    {4, 4, 6, chr_step_kernel<4, 4, 6>},
    {4, 4, 8, chr_step_kernel<4, 4, 8>},
};
}  // namespace

ChrKernelTable chr_table_g4() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
