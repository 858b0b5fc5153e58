// Instantiates the one-wavefront, coordinate-only CHR step kernel shapes that
// run four chains of one variant per workgroup with the variant's packed A in
// LDS (chain states in VGPRs, d <= 128), for the unpaired and the paired-row
// layouts, the latter also on scaled pairs; see chr_kernel.hip.
#include "chr_kernel_impl.h"

namespace naja {

namespace {
using chr_detail::chr_step_kernel;
const ChrKernelEntry kTable[] = {
    {1, 1, 1, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
     chr_step_kernel<1, 1, 1, true, 2, 2, false, kChrLdsChains>,
     chr_step_kernel<1, 1, 1, true, 2, 2, true, kChrLdsChains>,
     chr_step_kernel<1, 1, 1, true, 2, 2, true, kChrLdsChains, true>},
    {1, 1, 2, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
     chr_step_kernel<1, 1, 2, true, 2, 2, false, kChrLdsChains>,
     chr_step_kernel<1, 1, 2, true, 2, 2, true, kChrLdsChains>,
     chr_step_kernel<1, 1, 2, true, 2, 2, true, kChrLdsChains, true>},
    {1, 1, 3, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
     chr_step_kernel<1, 1, 3, true, 2, 2, false, kChrLdsChains>,
     chr_step_kernel<1, 1, 3, true, 2, 2, true, kChrLdsChains>,
     chr_step_kernel<1, 1, 3, true, 2, 2, true, kChrLdsChains, true>},
    {1, 1, 4, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
     chr_step_kernel<1, 1, 4, true, 2, 2, false, kChrLdsChains>,
     chr_step_kernel<1, 1, 4, true, 2, 2, true, kChrLdsChains>,
     chr_step_kernel<1, 1, 4, true, 2, 2, true, kChrLdsChains, true>},
    {1, 1, 5, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
     chr_step_kernel<1, 1, 5, true, 2, 2, false, kChrLdsChains>,
     chr_step_kernel<1, 1, 5, true, 2, 2, true, kChrLdsChains>,
     chr_step_kernel<1, 1, 5, true, 2, 2, true, kChrLdsChains, true>},
    {1, 1, 6, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
     chr_step_kernel<1, 1, 6, true, 2, 2, false, kChrLdsChains>,
     chr_step_kernel<1, 1, 6, true, 2, 2, true, kChrLdsChains>,
     chr_step_kernel<1, 1, 6, true, 2, 2, true, kChrLdsChains, true>},
    {1, 1, 7, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
     chr_step_kernel<1, 1, 7, true, 2, 2, false, kChrLdsChains>,
     chr_step_kernel<1, 1, 7, true, 2, 2, true, kChrLdsChains>,
     chr_step_kernel<1, 1, 7, true, 2, 2, true, kChrLdsChains, true>},
    {1, 1, 8, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
     chr_step_kernel<1, 1, 8, true, 2, 2, false, kChrLdsChains>,
     chr_step_kernel<1, 1, 8, true, 2, 2, true, kChrLdsChains>,
     chr_step_kernel<1, 1, 8, true, 2, 2, true, kChrLdsChains, true>},
};
}  // namespace

ChrKernelTable chr_table_w1l() { return ChrKernelTable{kTable, (int)(sizeof(kTable) / sizeof(kTable[0]))}; }

}  // namespace naja
