// lgca_tile_inst.hip -- instantiates the tiled step kernels of one (model, K)
// pair, or the batch kernel (all models) of one K.  The build compiles this
// file once per instantiation, with
//   -DLGCA_INST_KIND=<0 HPP | 1 FHP-I | 2 FHP-II/III | -1 batch> -DLGCA_INST_K=<1|2|4|8>,
// so they compile in parallel.
#include "lgca_tile.hpp"

#if !defined(LGCA_INST_KIND) || !defined(LGCA_INST_K)
#error "compile with -DLGCA_INST_KIND=<-1..2> -DLGCA_INST_K=<1|2|4|8>"
#endif

namespace lgca {

#if LGCA_INST_KIND >= 0
template hipError_t launch_tile_kk<LGCA_INST_KIND, LGCA_INST_K>(const TileArgs&, const TileOrder&, uint32_t,
                                                                bool, int, hipStream_t);
template int tile_waves_per_cu_kk<LGCA_INST_KIND, LGCA_INST_K>(bool);
#if LGCA_INST_K >= 4
template hipError_t launch_tile_pipe_kk<LGCA_INST_KIND, LGCA_INST_K>(const TileArgs&, const PipeArgs&, uint32_t,
                                                                     hipStream_t);
#endif
#else
template hipError_t launch_tile_batch_kk<LGCA_INST_K>(const BatchJob*, BatchState*, const WorkItem*, uint32_t,
                                                      uint32_t, bool, hipStream_t);
#endif

}  // namespace lgca
