// Instantiates CHR step kernel shapes (W wavefronts, G chains, R rows per
// lane) for the launch table; see chr_kernel.hip.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    {8, 1, 1, chr_step_kernel<8, 1, 1>},
    {8, 1, 2, chr_step_kernel<8, 1, 2>},
    {8, 1, 4, chr_step_kernel<8, 1, 4>},
    {8, 1, 6, chr_step_kernel<8, 1, 6>},
    {8, 1, 8, chr_step_kernel<8, 1, 8>},
    {8, 1, 10, chr_step_kernel<8, 1, 10>},
    {8, 1, 11, chr_step_kernel<8, 1, 11>},
    {8, 1, 12, chr_step_kernel<8, 1, 12>},
    {8, 1, 14, chr_step_kernel<8, 1, 14>},
    {8, 1, 16, chr_step_kernel<8, 1, 16>},
    {8, 1, 17, chr_step_kernel<8, 1, 17>},
    {8, 1, 18, chr_step_kernel<8, 1, 18>},
    {8, 1, 20, chr_step_kernel<8, 1, 20>},
    {8, 1, 22, chr_step_kernel<8, 1, 22>},
    {8, 1, 24, chr_step_kernel<8, 1, 24>},
};
}  // namespace

ChrKernelTable chr_table_w8() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
