// Instantiates the large-polytope CHR kernel (16 wavefronts, one chain per
// workgroup, R = 10..22 rows per lane, up to 22528 rows) for the launch table; see chr_kernel_impl.h.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_big_kernel;
const ChrKernelEntry kTable[] = {
    {16, 1, 10, chr_big_kernel<16, 10>}, {16, 1, 12, chr_big_kernel<16, 12>}, {16, 1, 14, chr_big_kernel<16, 14>},
    {16, 1, 16, chr_big_kernel<16, 16>}, {16, 1, 18, chr_big_kernel<16, 18>}, {16, 1, 20, chr_big_kernel<16, 20>},
    {16, 1, 21, chr_big_kernel<16, 21>}, {16, 1, 22, chr_big_kernel<16, 22>},
};
}  // namespace

ChrKernelTable chr_table_big() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
