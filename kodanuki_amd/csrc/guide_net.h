// Launchers of the learned guide's network kernels (guide_net.hip on the
// device; the CPU build of the library has its own of the same names).
#pragma once

#include "ctx.h"
#include "integrator_kernels.h"

namespace kdn {

// lobes of queries (kdn_guide_query each): 5 floats a lobe, and the
// selection probabilities (select may be null)
bool launchGuideInferQueries(const GuideView& g, const float* queries, uint32_t n, float* lobes, float* select, hipStream_t st);
// lobes of the guided vertices among the first n paths of a queue, into
// the per-path lobe buffer (lane-major)
bool launchGuideInferPaths(const GuideView& g, const PassParams& P, uint32_t queue, uint32_t n, float* lobes, hipStream_t st);
// forward and backward of samples [first, first + count) (modulo n):
// weight gradients summed into guideTrainPartials(count) partial sums of
// kGuideMlpParams floats, grid gradients added to gridGrad in fixed point
// (fixed_of; a wild one raises *flag), the loss to *loss; sample weights
// times *scale
uint32_t guideTrainPartials(uint32_t count);
bool launchGuideTrain(const GuideView& g, const float* samples, uint32_t n, uint32_t first, uint32_t count, const float* scale,
                      float* partial, uint64_t* gridGrad, float* loss, uint32_t* flag, hipStream_t st);

// a guided round (guide.hip): lobes and records before it, lobes of each
// guided bounce's vertices before their shading, training after its folds
bool guideBeginRound(Ctx& c, PassParams& P, uint32_t paths);
bool guideInferQueue(Ctx& c, const PassParams& P, uint32_t queue, uint32_t count);
bool guideEndRound(Ctx& c, const PassParams& P, uint32_t paths, kdn_pass_stats* stats);
// A round whose jobs have learned guides of their own (guides[j], null
// for a job without one; each such job has one plain slice a round):
// the jobs' learned/train_stride/path_first/rec_first fields and P's
// lobes and records set before it; the lobes of one job's stretch
// [first, first + count) of a queue; each guide trained on its own job's
// records after the folds.
bool guidesBeginRound(Ctx& c, PassParams& P, std::vector<JobDev>& jobs, const std::vector<SliceDev>& slices,
                      const std::vector<void*>& guides);
bool guidesInferQueue(Ctx& c, const PassParams& P, uint32_t queue, uint32_t first, uint32_t count, void* guide);
bool guidesEndRound(Ctx& c, const PassParams& P, const std::vector<JobDev>& jobs, const std::vector<SliceDev>& slices,
                    const std::vector<void*>& guides, kdn_pass_stats* stats);

}  // namespace kdn
