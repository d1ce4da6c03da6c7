// host/serve.h - a render-job server over the batched renderer: continuous batching with admission by path-state
// slots and scene memory, resident scenes with LRU eviction, and per-job latency.
//
// Jobs (scene, resolution, samples, bounce limit, seed, arrival time) are admitted in arrival order once they have
// arrived, their scene is resident (uploading it, and evicting the least recently used scenes no admitted job needs,
// within the scene budget) and the path-state slots their passes need are free. Every step runs one pass over all
// admitted jobs (each traces K_j more samples per pixel); a job whose samples are done leaves with its image, and the
// jobs that arrived meanwhile are admitted at the next step. The plain mode runs the same kernels one job at a time in
// arrival order. A job's image is the same in either mode and in any mix.
//
// The autotuner runs the submitted jobs under each combination of BVH width, LDS stack entries and rays per lane,
// keeps the fastest (rays per second over the whole mix) and leaves the jobs unfinished for the next run.
#pragma once
#include "device/batch.h"
#include "host/setup.h"
#include <memory>
#include <string>
#include <vector>

namespace cwp {

struct ServeOptions {
    BatchOptions batch;          // layout, triangle test, LDS stack, path-state slots, jobs per pass, bounce limit
    double sceneBudgetMB = 16384;// device memory for resident scenes
    bool batched = true;         // false: plain, one job at a time
};

struct ServeJob {
    int scene = 0, w = 0, h = 0, spp = 1, bounces = 6;
    uint32_t seed = 1000;
    double arrival = 0;          // seconds after the start of run()
    // filled by run()
    int K = 0, passes = 0;
    double admitted = -1, finished = -1;
    std::vector<float> image;    // per-pixel radiance sums (divide by spp), float4 per pixel
};

struct TuneTrial {
    int variant = 8, stackLds = 8, raysPerLane = 1;
    double raysPerS = 0, wallS = 0;
};

struct ServeStats {
    int jobs = 0, passes = 0, scenesLoaded = 0, scenesEvicted = 0;
    double wallS = 0, rays = 0, raysPerS = 0, latP50 = 0, latP99 = 0, latMean = 0, maxPathsInFlight = 0;
};

class Server {
public:
    bool init(const ServeOptions& O, std::string& err);
    // keeps a host copy of the scene (any BVH layout; others are built from its mesh when needed) and its camera /
    // lighting; uploaded when a job needs it
    int addScene(const SceneSetup& Z, std::string& err);
    int submit(const ServeJob& J, std::string& err);
    bool run(std::string& err);   // processes every submitted, unfinished job
    const ServeJob& job(int id) const { return jobs_[id]; }
    int numJobs() const { return (int)jobs_.size(); }
    size_t sceneBytes(int id) const;
    const ServeStats& stats() const { return stats_; }
    void clearJobs() { jobs_.clear(); }
    void resetJobs();             // marks every job unfinished (its image dropped)
    void setBatched(bool b) { O_.batched = b; }
    const ServeOptions& options() const { return O_; }
    // new BVH layout / LDS stack / rays per lane / graphs: the device renderer is rebuilt and every scene evicted
    bool reconfigure(const BatchOptions& B, std::string& err);
    // uploads the scenes of the unfinished jobs in arrival order while they fit the budget
    bool preload(std::string& err);
    bool autotune(const std::vector<int>& variants, const std::vector<int>& stacks, const std::vector<int>& raysPerLane,
                  std::vector<TuneTrial>& trials, std::string& err);
    // frame parameters of a job (the scene's camera and lighting at the job's resolution, bounces and seed)
    FrameParams jobParams(const ServeJob& J) const;
    // samples per pass of a job: all its samples if they fit a quarter of the path-state slots, else the largest
    // divisor of spp that fits
    int samplesPerPass(const ServeJob& J) const;

private:
    struct HostScene {
        std::shared_ptr<const SceneSetup> Z;   // mesh and builder options, for other layouts
        SceneUpload lay[3];  // layouts 2, 4, 8 (built when first needed)
        FrameParams P;       // at the setup's resolution
        float c2w[16];
        int w0 = 0, h0 = 0;
        int resident = -1;   // batch scene id when uploaded
        size_t bytes = 0;    // device bytes in the current layout
        double lastUse = 0;
    };
    static int layIndex(int variant) { return variant == 2 ? 0 : (variant == 4 ? 1 : 2); }
    bool ensureLayout(HostScene& H, std::string& err);
    bool makeResident(int s, const std::vector<int>& pinned, double now, std::string& err);
    ServeOptions O_;
    std::unique_ptr<BatchRenderer> R_;
    std::vector<HostScene> scenes_;
    std::vector<ServeJob> jobs_;
    ServeStats stats_;
    size_t residentBytes_ = 0;
};

}  // namespace cwp
