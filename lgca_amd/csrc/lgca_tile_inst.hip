// lgca_tile_inst.hip -- instantiates the tiled step kernels of one (model, K)
// pair.  The build compiles this file once per pair, with
//   -DLGCA_INST_KIND=<0 HPP | 1 FHP-I | 2 FHP-II/III> -DLGCA_INST_K=<1|2|4|8>,
// so the pairs compile in parallel.
#include "lgca_tile.hpp"

#if !defined(LGCA_INST_KIND) || !defined(LGCA_INST_K)
#error "compile with -DLGCA_INST_KIND=<0..2> -DLGCA_INST_K=<1|2|4|8>"
#endif

namespace lgca {

template hipError_t launch_tile_kk<LGCA_INST_KIND, LGCA_INST_K>(const TileArgs&, const TileOrder&, uint32_t,
                                                                bool, bool, hipStream_t);
template hipError_t launch_tile_batch_kk<LGCA_INST_KIND, LGCA_INST_K>(const BatchJob*, BatchState*,
                                                                      const WorkItem*, uint32_t, uint32_t,
                                                                      uint32_t, bool, hipStream_t);

}  // namespace lgca
