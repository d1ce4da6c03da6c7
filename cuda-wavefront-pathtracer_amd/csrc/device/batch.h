// device/batch.h - many render jobs in one wavefront pass (plain C++ interface; kernels in kernels.h).
//
// Scenes are uploaded once and stay resident until removed; a job is (scene, frame parameters, image, samples per
// pass). One pass traces K_j samples of every pixel of every job given: the jobs' paths are concatenated into one id
// space, so each bounce is one launch of the trace+shade kernel and one of the shadow kernel for the whole batch, and
// small jobs fill the GPU together. Each path's arithmetic, and each pixel's summation order, are those of the job
// rendered alone: a job's image does not depend on what it is batched with. The job table is copied to the device
// ahead of each pass and the kernels read the pass's sizes from it, so the pass's launches (counter reset, trace+shade
// and shadow per bounce, reduction, totals) are the same for every pass: they are captured once in a HIP graph per
// table slot and replayed.
#pragma once
#include "device/renderer.h"
#include <string>
#include <vector>

namespace cwp {

struct BatchOptions {
    int variant = 8;             // BVH layout of every scene (2, 4 or 8)
    int tri = kTriWoop;
    int stackLds = 8;            // 8, 16 or 32
    size_t maxPaths = 1 << 26;   // path-state slots (paths in flight per pass, all jobs together)
    int maxJobs = 256;           // jobs per pass
    int bounces = 8;             // deepest bounce limit of any job
    int raysPerLane = 1;         // 1, 2 or 4: the persistent grid is 1/raysPerLane of the resident waves, so each lane
                                 // traces that many times more rays in sequence
    int useGraph = 1;            // replay each pass's launches from a captured HIP graph
    int full = 0;                // 1: the full material integrator (every scene full); 0: the legacy integrator
    int compact = 0;             // queue compaction of rolling steps: 0 per wave (ballot, one atomic), 1 per lane
    int tail = -1;               // straggler hand-off (BVH8) of the trace+shade kernels, and with the full integrator
                                 // of the shadow kernels: a wave with fewer lanes than this still tracing hands their
                                 // traversals to a tail kernel; -1 kTailDefault with BVH8 (else off), 0 off, 2..64
    int tailLevels = 2;          // 2: the rolling steps' trace+shade tail kernels hand their own stragglers on to a
                                 // second tail (passes resume every record to the end)
};

struct BatchJob {
    int scene = -1;
    FrameParams F;               // camera (w, h), lighting, bounce limit, seed, filter
    int K = 1;                   // samples per pixel this pass
    int smpBase = 0;             // sample index of the pass's first sample
    void* accum = nullptr;       // device float4 per pixel (from allocImage)
};

// A live job of a rolling batch (see rollStep): its scene, parameters, samples per pixel of this round of samples
// (K, from sample index smpBase), image, and the path ids it owns, [pathBase, pathBase + w h K).
struct RollJob {
    int scene = -1;
    FrameParams F;
    int K = 1, smpBase = 0;
    void* accum = nullptr;
    int pathBase = 0;
};

class BatchRenderer {
public:
    BatchRenderer();
    ~BatchRenderer();
    BatchRenderer(const BatchRenderer&) = delete;
    BatchRenderer& operator=(const BatchRenderer&) = delete;

    bool init(const BatchOptions& O, std::string& err);
    // uploads a checked scene (same layout as the renderer's); returns its id, or -1 with err
    int addScene(const SceneUpload& S, std::string& err);
    void removeScene(int id);
    size_t sceneBytes(int id) const;
    // device bytes a scene takes when uploaded (nodes in the device layout, triangles, normals, albedo)
    static size_t uploadBytes(const SceneUpload& S, int tri);
    // per-pixel accumulation images (device, zeroed)
    bool allocImage(int nPix, void** img, std::string& err);
    void freeImage(void* img);
    bool clearImage(void* img, int nPix, std::string& err);
    bool readImage(const void* img, int nPix, float* rgba, std::string& err);
    // One pass over the given jobs, queued on the batch's stream; with `sync` it waits and st.frameMs is the pass's
    // device time. Ray counters accumulate over passes (counters()).
    bool pass(const std::vector<BatchJob>& jobs, bool sync, FrameStats& st, std::string& err);
    bool synchronize(std::string& err);
    // Rolling batches: steps instead of passes. Each step is one trace+shade launch and one shadow launch over every
    // live path of every live job (whatever its depth) plus the camera rays of the jobs admitted at that step, so a job
    // starts at the next launch rather than the next pass. A job admitted at step s has traced every ray at step
    // s + bounces + 1 and is reduced after that step (flag it done there); its path ids are free for the next step.
    // rollBegin clears the step counters; rollStep queues one step (no host synchronisation except when the step-table
    // ring wraps) and returns its id; stepFinished tells whether a step has completed on the device.
    bool rollBegin(std::string& err);
    bool rollStep(const std::vector<RollJob>& live, const std::vector<char>& isNew, const std::vector<char>& isDone,
                  long long& stepId, std::string& err);
    bool stepFinished(long long stepId) const;
    bool stepSlotFree() const;   // rollStep would not wait for an earlier step
    // an image copy on the batch's copy stream, which does not wait for queued steps (the caller makes sure the image
    // is final: the step that reduced it has finished); and a clear queued on the batch's stream, ahead of the step
    // that first writes the image
    bool fetchImage(const void* img, int nPix, float* rgba, std::string& err);
    bool clearImageAsync(void* img, int nPix, std::string& err);
    // ray counters summed over the passes since the last reset (waits for queued passes)
    bool counters(FrameStats& total, bool reset, std::string& err);
    const BatchOptions& options() const;
    double vramMB = 0;
    int gridBlocks = 0;

private:
    struct Impl;
    Impl* p_;
};

}  // namespace cwp
