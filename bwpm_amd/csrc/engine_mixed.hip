// Mixed instantiation of the engine: the two streaming planes of every slab step (X and the
// product planes) are stored as complex64 in HBM, the arithmetic and every other buffer are FP64.
#include "engine_impl.hpp"

namespace bwpm {

template class Engine<double, false, float>;

std::unique_ptr<EngineBase> make_engine_mixed(Spec spec) {
  return std::make_unique<Engine<double, false, float>>(std::move(spec));
}

void run_continuous_lane_mixed(LaneFeed& feed, int slots, int chunk, hipStream_t stream) {
  Engine<double, false, float>::continuous(feed, slots, chunk, stream);
}

}  // namespace bwpm
