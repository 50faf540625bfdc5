// Block-floating-point instantiation of the engine: the two streaming planes of every slab step
// (X and the product planes) are stored in the 8-byte block-floating-point complex format (bfp,
// kernels.hpp), the size of complex64 with a rounding error 8 to 16 times smaller; the arithmetic
// and every other buffer are FP64.
#include "engine_impl.hpp"

namespace bwpm {

template class Engine<double, false, bfp>;

std::unique_ptr<EngineBase> make_engine_bfp(Spec spec) {
  return std::make_unique<Engine<double, false, bfp>>(std::move(spec));
}

void run_continuous_lane_bfp(LaneFeed& feed, int slots, int chunk, hipStream_t stream) {
  Engine<double, false, bfp>::continuous(feed, slots, chunk, stream);
}

}  // namespace bwpm
