// host/serve.cpp - see serve.h.
#include "host/serve.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>

namespace cwp {

namespace {
double nowS(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const double x = q * (double)(v.size() - 1);
    const size_t i = (size_t)x;
    const double f = x - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}
}  // namespace

bool Server::init(const ServeOptions& O, std::string& err) {
    if (O.sceneBudgetMB <= 0) { err = "serve: scene budget must be positive"; return false; }
    O_ = O;
    return reconfigure(O.batch, err);
}

bool Server::reconfigure(const BatchOptions& B, std::string& err) {
    for (HostScene& H : scenes_) { H.resident = -1; H.bytes = 0; }
    residentBytes_ = 0;
    R_.reset();   // frees the scenes and path state of the previous configuration
    std::unique_ptr<BatchRenderer> R(new BatchRenderer());
    if (!R->init(B, err)) return false;
    R_ = std::move(R);
    O_.batch = R_->options();   // as resolved by the renderer (the automatic hand-off threshold)
    return true;
}

size_t Server::sceneBytes(int id) const {
    if (id < 0 || id >= (int)scenes_.size()) return 0;
    const SceneUpload& S = scenes_[id].lay[layIndex(O_.batch.variant)];
    return S.nodes.empty() ? 0 : BatchRenderer::uploadBytes(S, O_.batch.tri);
}

bool Server::ensureLayout(HostScene& H, std::string& err) {
    SceneUpload& S = H.lay[layIndex(O_.batch.variant)];
    if (S.nodes.empty() && !relayoutScene(*H.Z, O_.batch.variant, S, err)) return false;
    H.bytes = BatchRenderer::uploadBytes(S, O_.batch.tri);
    return true;
}

int Server::addScene(const SceneSetup& Z, std::string& err) {
    if (Z.inst) { err = "serve: two-level scenes are not served"; return -1; }
    if (!nodeFloat4s(Z.variant)) { err = "serve: the setup has no BVH"; return -1; }
    if (Z.full && !O_.batch.full) { err = "serve: the scene needs the full integrator; the server runs the legacy one"; return -1; }
    HostScene H;
    std::shared_ptr<SceneSetup> Zc = std::make_shared<SceneSetup>(Z);
    // a legacy setup on a full-integrator server renders its Lambert materials with the full integrator: its layout
    // is rebuilt with the full-integrator tables
    if (O_.batch.full && !Z.full) Zc->full = true;
    else H.lay[layIndex(Z.variant)] = std::move(Zc->S);
    Zc->S = SceneUpload(); Zc->tree = BinaryBVH();   // the copy keeps the mesh and options only
    H.Z = Zc;
    H.P = Z.P;
    for (int k = 0; k < 16; k++) H.c2w[k] = Z.c2w[k];
    H.w0 = Z.A.w; H.h0 = Z.A.h;
    if (!ensureLayout(H, err)) return -1;
    scenes_.push_back(std::move(H));
    return (int)scenes_.size() - 1;
}

void Server::resetJobs() {
    for (ServeJob& J : jobs_) { J.passes = 0; J.admitted = -1; J.finished = -1; J.image.clear(); }
}

FrameParams Server::jobParams(const ServeJob& J) const {
    const HostScene& H = scenes_[J.scene];
    FrameParams F = H.P;   // camera, lighting, filter and epsilon of the setup
    const float s = (float)J.w / (float)H.w0;   // the same field of view at the job's width
    F.cam = makeCamera(H.c2w, H.P.cam.fx * s, H.P.cam.fy * s, J.w, J.h);
    F.cam.lensR = H.P.cam.lensR; F.cam.focusD = H.P.cam.focusD;
    F.maxB = J.bounces; F.seed = J.seed;
    return F;
}

int Server::samplesPerPass(const ServeJob& J) const {
    const size_t N = (size_t)J.w * J.h, cap = std::max<size_t>(1, O_.batch.maxPaths / 4);
    for (int K = J.spp; K > 1; K--)
        if (J.spp % K == 0 && N * (size_t)K <= cap) return K;
    return 1;
}

int Server::submit(const ServeJob& J, std::string& err) {
    if (J.scene < 0 || J.scene >= (int)scenes_.size()) { err = "serve: unknown scene"; return -1; }
    if (J.w < 1 || J.h < 1 || J.spp < 1 || J.bounces < 0 || J.bounces > O_.batch.bounces || J.arrival < 0) {
        err = "serve: bad job (size, samples, bounces or arrival)"; return -1;
    }
    if ((size_t)J.w * J.h > O_.batch.maxPaths) { err = "serve: the job's image exceeds the path-state slots"; return -1; }
    ServeJob j = J;
    j.K = samplesPerPass(J); j.passes = 0; j.admitted = -1; j.finished = -1; j.image.clear();
    jobs_.push_back(j);
    return (int)jobs_.size() - 1;
}

// uploads scene s if needed, evicting least recently used scenes that no job in `pinned` uses while over budget
bool Server::makeResident(int s, const std::vector<int>& pinned, double now, std::string& err) {
    HostScene& H = scenes_[s];
    H.lastUse = now;
    if (H.resident >= 0) return true;
    if (!ensureLayout(H, err)) return false;
    const size_t budget = (size_t)(O_.sceneBudgetMB * 1048576.0);
    while (residentBytes_ + H.bytes > budget) {
        int victim = -1;
        for (int k = 0; k < (int)scenes_.size(); k++) {
            if (scenes_[k].resident < 0 || std::find(pinned.begin(), pinned.end(), k) != pinned.end()) continue;
            if (victim < 0 || scenes_[k].lastUse < scenes_[victim].lastUse) victim = k;
        }
        if (victim < 0) return false;   // everything resident is in use: wait
        R_->removeScene(scenes_[victim].resident);
        scenes_[victim].resident = -1;
        residentBytes_ -= scenes_[victim].bytes;
        stats_.scenesEvicted++;
    }
    const int id = R_->addScene(H.lay[layIndex(O_.batch.variant)], err);
    if (id < 0) return false;
    H.resident = id;
    residentBytes_ += H.bytes;
    stats_.scenesLoaded++;
    return true;
}

bool Server::preload(std::string& err) {
    err.clear();
    if (!R_) { err = "serve: no device renderer (a reconfiguration failed)"; return false; }
    std::vector<int> order;
    for (int j = 0; j < (int)jobs_.size(); j++) if (jobs_[j].finished < 0) order.push_back(j);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return jobs_[a].arrival < jobs_[b].arrival; });
    std::vector<int> want;
    for (int j : order) if (std::find(want.begin(), want.end(), jobs_[j].scene) == want.end()) want.push_back(jobs_[j].scene);
    const size_t budget = (size_t)(O_.sceneBudgetMB * 1048576.0);
    for (int s : want) {
        if (!ensureLayout(scenes_[s], err)) return false;
        if (scenes_[s].resident < 0 && residentBytes_ + scenes_[s].bytes > budget) break;
        if (!makeResident(s, want, 0.0, err)) return false;
    }
    return true;
}

bool Server::autotune(const std::vector<int>& variants, const std::vector<int>& stacks, const std::vector<int>& raysPerLane,
                      const std::vector<int>& compacts, std::vector<TuneTrial>& trials, std::string& err) {
    trials.clear();
    if (running()) { err = "autotune: the live worker is running"; return false; }
    if (variants.empty() || stacks.empty() || raysPerLane.empty() || compacts.empty()) { err = "autotune: empty search space"; return false; }
    const bool batched = O_.batched;
    O_.batched = true;
    const BatchOptions base = O_.batch;
    TuneTrial best;
    best.raysPerS = -1;
    for (int v : variants)
        for (int sl : stacks)
            for (int r : raysPerLane)
                for (int cm : compacts) {
                    BatchOptions B = base;
                    B.variant = v; B.stackLds = sl; B.raysPerLane = r; B.compact = cm;
                    if (v != 8) B.tail = 0;   // the hand-off resumes BVH8 traversals only
                    resetJobs();
                    if (!reconfigure(B, err) || !preload(err) || !run(err)) { O_.batched = batched; return false; }
                    TuneTrial t;
                    t.variant = v; t.stackLds = sl; t.raysPerLane = r; t.compact = cm;
                    t.raysPerS = stats_.raysPerS; t.wallS = stats_.wallS;
                    trials.push_back(t);
                    if (t.raysPerS > best.raysPerS) best = t;
                }
    O_.batched = batched;
    BatchOptions B = base;
    B.variant = best.variant; B.stackLds = best.stackLds; B.raysPerLane = best.raysPerLane; B.compact = best.compact;
    if (B.variant != 8) B.tail = 0;
    resetJobs();
    stats_ = ServeStats();
    return reconfigure(B, err);
}

bool Server::run(std::string& err) {
    err.clear();
    if (running()) { err = "serve: the live worker is running (use submitLive / stop)"; return false; }
    if (!R_) { err = "serve: no device renderer (a reconfiguration failed)"; return false; }
    if (O_.batched && O_.rolling) return runRolling(false, err);
    return runPasses(err);
}

// Rolling continuous batching (see serve.h). Path ids are handed out as contiguous ranges (first fit in a sorted free
// list); images come from a pool (no device allocation or free while steps are queued, after the first use of a
// size); the host builds and queues a step whenever the step-table ring has room, and polls finished steps for the
// images of the jobs they completed.
bool Server::runRolling(bool live, std::string& err) {
    std::unique_lock<std::mutex> lock(mu_, std::defer_lock);
    if (live) lock.lock();
    stats_ = ServeStats();
    FrameStats c;
    if (!R_->counters(c, true, err) || !R_->rollBegin(err)) return false;
    const auto t0 = live ? liveT0_ : std::chrono::steady_clock::now();
    std::vector<std::pair<size_t, size_t>> freeRanges{ { 0, O_.batch.maxPaths } };   // [start, end), ascending
    auto takeRange = [&](size_t need, int& base) {
        for (size_t k = 0; k < freeRanges.size(); k++) {
            if (freeRanges[k].second - freeRanges[k].first < need) continue;
            base = (int)freeRanges[k].first;
            freeRanges[k].first += need;
            if (freeRanges[k].first == freeRanges[k].second) freeRanges.erase(freeRanges.begin() + (long)k);
            return true;
        }
        return false;
    };
    auto giveRange = [&](size_t b, size_t n) {
        auto it = std::lower_bound(freeRanges.begin(), freeRanges.end(), std::make_pair(b, b));
        it = freeRanges.insert(it, { b, b + n });
        if (it + 1 != freeRanges.end() && it->second == (it + 1)->first) { it->second = (it + 1)->second; freeRanges.erase(it + 1); }
        if (it != freeRanges.begin() && (it - 1)->second == it->first) { (it - 1)->second = it->second; freeRanges.erase(it); }
    };
    struct Img { void* p; int cap; };
    std::vector<Img> pool;                       // free images
    std::vector<std::pair<int, Img>> held;       // job -> its image while it runs or waits for its read-back
    auto imageOf = [&](int job) -> Img* { for (auto& h : held) if (h.first == job) return &h.second; return nullptr; };
    struct LiveJob { int job, base; size_t paths; long long first, last; };
    std::vector<LiveJob> liveJobs;
    struct Waiting { int job; long long step; };
    std::vector<Waiting> waiting;                // done jobs whose last step has not finished yet
    std::vector<double> lat;
    std::vector<char> busy;                      // per job: live or waiting
    long long step = 0;
    bool ok = true;
    auto freePool = [&]() {
        std::string e;
        R_->synchronize(e);
        for (Img& i : pool) R_->freeImage(i.p);
        for (auto& h : held) R_->freeImage(h.second.p);
        pool.clear(); held.clear();
    };
    for (;;) {
        const double now = nowS(t0);
        // 1. read back the images of jobs whose last step has finished
        for (size_t k = 0; k < waiting.size();) {
            if (!R_->stepFinished(waiting[k].step)) { k++; continue; }
            const int j = waiting[k].job;
            ServeJob& J = jobs_[j];
            Img* im = imageOf(j);
            J.image.resize((size_t)J.w * J.h * 4);
            if (!R_->fetchImage(im->p, J.w * J.h, J.image.data(), err)) { ok = false; break; }
            pool.push_back(*im);
            held.erase(std::find_if(held.begin(), held.end(), [&](const std::pair<int, Img>& h) { return h.first == j; }));
            J.finished = nowS(t0);
            lat.push_back(J.finished - J.arrival);
            stats_.jobs++;
            busy[j] = 0;
            waiting.erase(waiting.begin() + (long)k);
            if (live) cv_.notify_all();
        }
        if (!ok) break;
        // 2. admission: arrived, unfinished, idle jobs in arrival order, first fit; a job that has waited longer than
        //    the starvation limit stops later ones from overtaking it
        if (busy.size() < jobs_.size()) busy.resize(jobs_.size(), 0);
        std::vector<int> cand;
        double nextArrival = 1e300;
        bool unfinished = false;
        for (int j = 0; j < (int)jobs_.size(); j++) {
            const ServeJob& J = jobs_[j];
            if (J.finished >= 0) continue;
            unfinished = true;
            if (busy[j]) continue;
            if (J.arrival > now) { nextArrival = std::min(nextArrival, J.arrival); continue; }
            cand.push_back(j);
        }
        std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) { return jobs_[a].arrival < jobs_[b].arrival; });
        for (int j : cand) {
            if ((int)liveJobs.size() >= O_.batch.maxJobs) break;
            ServeJob& J = jobs_[j];
            const bool starving = now - J.arrival > O_.starveS;
            const size_t need = (size_t)J.w * J.h * J.K;
            int base = 0;
            if (!takeRange(need, base)) { if (starving) break; continue; }
            std::vector<int> pinned;
            for (const LiveJob& L : liveJobs) pinned.push_back(jobs_[L.job].scene);
            pinned.push_back(J.scene);
            if (!makeResident(J.scene, pinned, now, err)) {
                giveRange((size_t)base, need);
                if (!err.empty()) { ok = false; break; }
                if (liveJobs.empty() && waiting.empty()) { err = "serve: a job's scene does not fit the scene budget"; ok = false; break; }
                if (starving) break;
                continue;
            }
            if (J.passes == 0) {   // first round: an image from the pool (or a new one), cleared ahead of its first step
                Img im{ nullptr, 0 };
                for (size_t k = 0; k < pool.size(); k++)
                    if (pool[k].cap >= J.w * J.h) { im = pool[k]; pool.erase(pool.begin() + (long)k); break; }
                if (!im.p) {
                    im.cap = J.w * J.h;
                    if (!R_->allocImage(im.cap, &im.p, err)) { ok = false; break; }
                } else if (!R_->clearImageAsync(im.p, J.w * J.h, err)) { ok = false; break; }
                held.push_back({ j, im });
                J.admitted = now;
            }
            busy[j] = 1;
            liveJobs.push_back({ j, base, need, step, step + J.bounces + 1 });
            stats_.passes++;
            scenes_[J.scene].lastUse = now;
        }
        if (!ok) break;
        if (liveJobs.empty()) {
            if (waiting.empty() && !unfinished) {
                if (!live || stopReq_) break;
                cv_.wait_for(lock, std::chrono::milliseconds(1));   // a live worker idles until a submission
                continue;
            }
            const double wait = std::min(nextArrival - nowS(t0), 2e-4);
            if (live) lock.unlock();
            std::this_thread::sleep_for(std::chrono::duration<double>(std::max(wait, 2e-5)));
            if (live) lock.lock();
            continue;
        }
        // 3. the next step, once the step-table ring has room (never block on the device while jobs may arrive)
        if (!R_->stepSlotFree()) {
            if (live) lock.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(50));
            if (live) lock.lock();
            continue;
        }
        std::vector<RollJob> rj;
        std::vector<char> isNew, isDone;
        for (const LiveJob& L : liveJobs) {
            const ServeJob& J = jobs_[L.job];
            RollJob r;
            r.scene = scenes_[J.scene].resident; r.F = jobParams(J); r.K = J.K; r.smpBase = J.passes * J.K;
            r.accum = imageOf(L.job)->p; r.pathBase = L.base;
            rj.push_back(r); isNew.push_back(L.first == step); isDone.push_back(L.last == step);
        }
        long long id = 0;
        if (!R_->rollStep(rj, isNew, isDone, id, err)) { ok = false; break; }
        stats_.steps++;
        stats_.maxPathsInFlight = std::max(stats_.maxPathsInFlight, (double)(O_.batch.maxPaths - [&] {
            size_t f = 0; for (auto& r : freeRanges) f += r.second - r.first; return f; }()));
        // jobs whose round ended at this step: their path ids are free from the next step on
        std::vector<LiveJob> keep;
        for (const LiveJob& L : liveJobs) {
            if (L.last != step) { keep.push_back(L); continue; }
            giveRange((size_t)L.base, L.paths);
            ServeJob& J = jobs_[L.job];
            J.passes++;
            if (J.passes * J.K >= J.spp) waiting.push_back({ L.job, id });
            else busy[L.job] = 0;   // its next round is admitted like an arrival
        }
        liveJobs.swap(keep);
        step++;
    }
    freePool();
    if (!ok) return false;
    stats_.wallS = nowS(t0);
    if (!R_->counters(c, true, err)) return false;
    stats_.rays = c.rays + c.shadowRays;
    stats_.raysPerS = stats_.wallS > 0 ? stats_.rays / stats_.wallS : 0;
    stats_.latP50 = percentile(lat, 0.5); stats_.latP99 = percentile(lat, 0.99);
    double sum = 0; for (double x : lat) sum += x;
    stats_.latMean = lat.empty() ? 0 : sum / (double)lat.size();
    return true;
}

bool Server::start(std::string& err) {
    if (running()) { err = "serve: already started"; return false; }
    if (!R_) { err = "serve: no device renderer (a reconfiguration failed)"; return false; }
    {
        std::lock_guard<std::mutex> lk(mu_);
        stopReq_ = false; workerOk_ = true; workerErr_.clear();
        liveT0_ = std::chrono::steady_clock::now();
    }
    worker_ = std::thread([this] {
        std::string e;
        const bool ok = runRolling(true, e);
        std::lock_guard<std::mutex> lk(mu_);
        workerOk_ = ok; workerErr_ = e;
        cv_.notify_all();
    });
    return true;
}

int Server::submitLive(ServeJob J, std::string& err) {
    if (!running()) { err = "serve: not started"; return -1; }
    std::lock_guard<std::mutex> lk(mu_);
    if (stopReq_) { err = "serve: stopping"; return -1; }
    J.arrival = std::chrono::duration<double>(std::chrono::steady_clock::now() - liveT0_).count();
    if (J.scene < 0 || J.scene >= (int)scenes_.size()) { err = "serve: unknown scene"; return -1; }
    if (J.w < 1 || J.h < 1 || J.spp < 1 || J.bounces < 0 || J.bounces > O_.batch.bounces) {
        err = "serve: bad job (size, samples or bounces)"; return -1;
    }
    if ((size_t)J.w * J.h > O_.batch.maxPaths) { err = "serve: the job's image exceeds the path-state slots"; return -1; }
    J.K = samplesPerPass(J); J.passes = 0; J.admitted = -1; J.finished = -1; J.image.clear();
    jobs_.push_back(J);
    cv_.notify_all();
    return (int)jobs_.size() - 1;
}

JobState Server::poll(int id) const {
    std::lock_guard<std::mutex> lk(mu_);
    if (id < 0 || id >= (int)jobs_.size()) return kJobQueued;
    const ServeJob& J = jobs_[id];
    return J.finished >= 0 ? kJobDone : (J.admitted >= 0 ? kJobRunning : kJobQueued);
}

bool Server::result(int id, std::vector<float>& rgba, std::string& err) const {
    std::lock_guard<std::mutex> lk(mu_);
    if (id < 0 || id >= (int)jobs_.size()) { err = "serve: unknown job"; return false; }
    const ServeJob& J = jobs_[id];
    if (J.finished < 0) { err = "serve: the job has not finished"; return false; }
    rgba = J.image;
    return true;
}

bool Server::stop(std::string& err) {
    if (!running()) { err = "serve: not started"; return false; }
    {
        std::lock_guard<std::mutex> lk(mu_);
        stopReq_ = true;
        cv_.notify_all();
    }
    worker_.join();
    std::lock_guard<std::mutex> lk(mu_);
    if (!workerOk_) err = workerErr_;
    return workerOk_;
}

Server::~Server() {
    if (running()) { std::string e; stop(e); }
}

bool Server::runPasses(std::string& err) {
    stats_ = ServeStats();
    FrameStats c;
    if (!R_->counters(c, true, err)) return false;
    std::vector<int> order;   // unfinished jobs in arrival order
    for (int j = 0; j < (int)jobs_.size(); j++) if (jobs_[j].finished < 0) order.push_back(j);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return jobs_[a].arrival < jobs_[b].arrival; });
    struct Active { int job; void* img; };
    std::vector<Active> active;
    size_t next = 0, pathsInUse = 0;
    std::vector<double> lat;
    const auto t0 = std::chrono::steady_clock::now();
    const size_t slots = O_.batch.maxPaths;
    while (next < order.size() || !active.empty()) {
        double now = nowS(t0);
        // admission, in arrival order (a job that does not fit blocks the ones behind it: no starvation)
        while (next < order.size() && (O_.batched || active.empty())) {
            ServeJob& J = jobs_[order[next]];
            if (J.arrival > now) break;
            if ((int)active.size() >= O_.batch.maxJobs) break;
            const size_t need = (size_t)J.w * J.h * J.K;
            if (pathsInUse + need > slots) break;
            std::vector<int> pinned;
            for (const Active& a : active) pinned.push_back(jobs_[a.job].scene);
            pinned.push_back(J.scene);
            if (!makeResident(J.scene, pinned, now, err)) {
                if (!err.empty()) return false;
                if (active.empty()) { err = "serve: a job's scene does not fit the scene budget"; return false; }
                break;
            }
            void* img = nullptr;
            if (!R_->allocImage(J.w * J.h, &img, err)) return false;
            J.admitted = now;
            active.push_back({ order[next], img });
            pathsInUse += need;
            next++;
        }
        if (active.empty()) {   // idle until the next arrival
            const double wait = jobs_[order[next]].arrival - nowS(t0);
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.001)));
            continue;
        }
        stats_.maxPathsInFlight = std::max(stats_.maxPathsInFlight, (double)pathsInUse);
        std::vector<BatchJob> pass;
        for (const Active& a : active) {
            const ServeJob& J = jobs_[a.job];
            BatchJob B;
            B.scene = scenes_[J.scene].resident; B.F = jobParams(J); B.K = J.K; B.smpBase = J.passes * J.K; B.accum = a.img;
            pass.push_back(B);
            scenes_[J.scene].lastUse = now;
        }
        FrameStats st;
        if (!R_->pass(pass, true, st, err)) return false;
        stats_.passes++;
        now = nowS(t0);
        std::vector<Active> keep;
        for (const Active& a : active) {
            ServeJob& J = jobs_[a.job];
            J.passes++;
            if (J.passes * J.K < J.spp) { keep.push_back(a); continue; }
            J.image.resize((size_t)J.w * J.h * 4);
            if (!R_->readImage(a.img, J.w * J.h, J.image.data(), err)) return false;
            R_->freeImage(a.img);
            J.finished = nowS(t0);
            lat.push_back(J.finished - J.arrival);
            pathsInUse -= (size_t)J.w * J.h * J.K;
            stats_.jobs++;
        }
        active.swap(keep);
    }
    stats_.wallS = nowS(t0);
    if (!R_->counters(c, true, err)) return false;
    stats_.rays = c.rays + c.shadowRays;
    stats_.raysPerS = stats_.wallS > 0 ? stats_.rays / stats_.wallS : 0;
    stats_.latP50 = percentile(lat, 0.5); stats_.latP99 = percentile(lat, 0.99);
    double sum = 0; for (double x : lat) sum += x;
    stats_.latMean = lat.empty() ? 0 : sum / (double)lat.size();
    return true;
}

}  // namespace cwp
