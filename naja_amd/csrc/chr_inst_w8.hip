// Instantiates CHR step kernel shapes (W wavefronts, G chains, R rows per
// lane) for the launch table; see chr_kernel.hip.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    {8, 1, 17, chr_step_kernel<8, 1, 17>},
    {8, 1, 18, chr_step_kernel<8, 1, 18>},
    {8, 1, 20, chr_step_kernel<8, 1, 20>},
    {8, 1, 22, chr_step_kernel<8, 1, 22>},
    {8, 1, 24, chr_step_kernel<8, 1, 24>},
    {8, 1, 26, chr_step_kernel<8, 1, 26>},
    {8, 1, 28, chr_step_kernel<8, 1, 28>},
    {8, 1, 30, chr_step_kernel<8, 1, 30>},
    {8, 1, 32, chr_step_kernel<8, 1, 32>},
};
}  // namespace

ChrKernelTable chr_table_w8() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
