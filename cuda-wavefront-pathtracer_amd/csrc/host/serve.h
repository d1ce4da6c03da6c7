// host/serve.h - a render-job server over the batched renderer: continuous batching with admission by path-state
// slots and scene memory, resident scenes with LRU eviction, and per-job latency.
//
// Jobs (scene, resolution, samples, bounce limit, seed, arrival time) are admitted once they have arrived, their scene
// is resident (uploading it, and evicting the least recently used scenes no live job needs, within the scene budget)
// and a contiguous range of path ids for K_j samples of every pixel is free. Batched mode is ROLLING (continuous
// batching at bounce granularity): the device runs steps, each one trace+shade and one shadow launch over every live
// path of every live job plus the camera rays of the jobs admitted at that step, so an arrival joins the very next
// launch; admission is first fit in arrival order (a job that does not fit yet does not block smaller ones behind it,
// and a job that waits longer than the starvation limit stops later arrivals from overtaking it); the host never waits
// for a step, it only polls which steps have finished. A job whose samples are done leaves with its image and frees
// its path ids for the next step. The pass mode (rolling off) admits between whole multi-bounce passes; the plain mode
// runs the same kernels one job at a time in arrival order. A job's image is the same in every mode and in any mix.
// start() runs the rolling loop on a worker thread: submit() / poll() / result() may then be called from other threads
// while jobs are on the GPU; stop() lets the queued jobs finish and joins the worker.
//
// The autotuner runs the submitted jobs under each combination of BVH width, LDS stack entries, rays per lane and
// queue-compaction granularity (per wave or per lane), keeps the fastest (rays per second over the whole mix) and
// leaves the jobs unfinished for the next run.
#pragma once
#include "device/batch.h"
#include "host/setup.h"
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace cwp {

struct ServeOptions {
    BatchOptions batch;          // layout, triangle test, LDS stack, path-state slots, jobs per pass, bounce limit
    double sceneBudgetMB = 16384;// device memory for resident scenes
    bool batched = true;         // false: plain, one job at a time
    bool rolling = true;         // batched mode: admission at every step (true) or between whole passes (false)
    double starveS = 0.25;       // rolling admission: a job waiting longer than this stops later jobs from overtaking it
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
    int variant = 8, stackLds = 8, raysPerLane = 1, compact = 0;
    double raysPerS = 0, wallS = 0;
};

// job states as poll() reports them
enum JobState : int { kJobQueued = 0, kJobRunning = 1, kJobDone = 2 };

struct ServeStats {
    int jobs = 0, passes = 0, scenesLoaded = 0, scenesEvicted = 0;
    long long steps = 0;         // rolling steps (each one trace+shade and one shadow launch)
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
    // live serving: the rolling loop on a worker thread; submitLive stamps the arrival with the current time
    bool start(std::string& err);
    int submitLive(ServeJob J, std::string& err);
    JobState poll(int id) const;
    bool result(int id, std::vector<float>& rgba, std::string& err) const;
    bool stop(std::string& err);  // waits until every submitted job is done, then joins the worker
    bool running() const { return worker_.joinable(); }
    ~Server();
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
                  const std::vector<int>& compacts, std::vector<TuneTrial>& trials, std::string& err);
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
    bool runPasses(std::string& err);
    bool runRolling(bool live, std::string& err);
    bool ensureLayout(HostScene& H, std::string& err);
    bool makeResident(int s, const std::vector<int>& pinned, double now, std::string& err);
    ServeOptions O_;
    std::unique_ptr<BatchRenderer> R_;
    std::vector<HostScene> scenes_;
    std::deque<ServeJob> jobs_;   // a deque: live submissions append without moving the jobs the worker holds
    ServeStats stats_;
    size_t residentBytes_ = 0;
    // live serving
    mutable std::mutex mu_;
    std::condition_variable cv_;
    std::thread worker_;
    bool stopReq_ = false, workerOk_ = true;
    std::string workerErr_;
    std::chrono::steady_clock::time_point liveT0_;
};

}  // namespace cwp
