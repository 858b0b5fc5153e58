// Instantiates CHR step kernel shapes (W wavefronts, G chains, R rows per
// lane) for the launch table; see chr_kernel.hip.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    {16, 1, 1, chr_step_kernel<16, 1, 1>},
    {16, 1, 2, chr_step_kernel<16, 1, 2>},
    {16, 1, 4, chr_step_kernel<16, 1, 4>},
    {16, 1, 6, chr_step_kernel<16, 1, 6>},
    {16, 1, 8, chr_step_kernel<16, 1, 8>},
};
}  // namespace

ChrKernelTable chr_table_w16() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
