// batch.hip -- many clips of uneven sizes (frame size, grid, window, search range, frame count) in one
// hipGraph launch
//
// A batch holds one correlator per clip (correlator.h: its grid, reference windows and box-sum
// buffers). Clips with the same tile shape -- the int8 form: the same window, search range and SNR
// box side min(stat_size, M) (one kernel instantiation and LDS layout); the FFT form: the same
// search-window size L -- share a persistent kernel that drains a work queue of (clip, window, run of
// frames) items across all of them (WorkQueue in common.h: one range per XCD, then stealing), so small
// clips fill the GPU together and large ones spread over every CU. A call is: zero the queue counters (a
// kernel, not a memset node), then per stage of frames (each clip's frames in groups its box-sum buffers
// hold) the box-sum maps and one queue kernel per tile shape. The int8 form runs this as one chain: the
// maps of every clip in one launch, then the shapes' kernels one after another (each fills every
// register slot of the CUs at its tuned blocks per CU; run side by side they slowed each other, 6 %).
// The FFT form gives every tile shape a branch stream of its own (at most kBranches, the shapes dealt
// largest first to the least loaded): per stage the maps of the shape's clips and its kernel. The
// shapes share nothing but the read-only argument table (each clip has its own maps, counters and
// outputs), and the FFT kernels' one block per CU leaves room beside it, so the branches run side by
// side (12 % less time on the 16-clip workload). That sequence is captured once into a hipGraph and
// replayed while the buffers stay the same; the clips' arguments sit in a device table written before
// the capture.
//
// The autotuner: per tile shape, on the first call (when enabled), every candidate (int8: kernel variant
// x frames per item x resident blocks per CU; FFT: frames per item x blocks per CU) is timed with
// hipEvents on a sample of that shape's frames, and the fastest is kept. Every candidate computes the
// same results: the variants give the same offsets, peaks and covariances and the SNR to rounding, the
// item size and the grid only move work between blocks.
//
// No counterpart in the reference (one clip per correlator, one launch sequence per group of frames,
// cxx/cuda.cu:317-360).

#include "correlator.h"

using namespace framecor;

namespace {

__global__ void zeroInts(int *p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

struct Shape {
    int W, R, L;
    int snr;            // the int8 form: the SNR box side min(stat, M), which sizes the kernel's LDS
    std::vector<int> clips;
    Int8Layout layout{};
    // the plan
    int variant = int(Int8Variant::Auto);
    int fpb = 0;        // frames per item; 0: automatic
    int perCu = 0;      // resident blocks per CU; 0: as many as fit
    bool tuned = false;
    // the FFT form with spectra beyond LDS: the per-block workspace of at most workBlocks resident blocks
    std::shared_ptr<DeviceBuffer> work;
    int workBlocks = 0;
    // the launch of the plan (prepared before a capture: no allocation or query inside one)
    const void *func = nullptr;
    int threads = 0;
    size_t shmem = 0;
    int blocks = 0;
};

// the pointers of one call (a new set means a new capture)
struct CallPointers {
    std::vector<const void *> p;
    bool operator==(const CallPointers &o) const { return p == o.p; }
};

constexpr int kBranches = 4;   // the branch streams of a call (the hardware queues a process gets by default)

} // namespace

struct framecor_batch {
    int method = FRAMECOR_METHOD_INT8;
    int device = 0;
    int cus = 0;
    std::vector<std::unique_ptr<framecor_correlator>> clips;
    std::vector<int> frames;        // frames of each clip per call
    std::vector<Shape> shapes;
    std::vector<int> shapeOf;       // clip -> shape
    int stages = 0;
    bool tune = false;
    hipStream_t stream = nullptr;   // capture and tuning
    hipEvent_t lastLaunch = nullptr;
    bool launched = false;
    DeviceBuffer dArgs, dCounters;
    // the box-map launches of each stage (boxGroups of them: the int8 form one for every clip, the FFT
    // form one per tile shape), index stage * boxGroups + group: jobCount jobs from jobFirst in dJobs,
    // jobBlocks blocks
    DeviceBuffer dJobs;
    std::vector<int> jobFirst, jobCount, jobBlocks;
    int boxGroups = 1;
    // the FFT form: the branch streams of a call and the events that fork them from `stream` and join
    // them to it; branchOf[shape] its branch
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> joined;
    hipEvent_t forked = nullptr;
    std::vector<int> branchOf;
    // the pointers of the call whose arguments are in dArgs (the autotuner rewrites entries of stage 0)
    const uint8_t *const *argFr = nullptr;
    const int32_t *const *argGross = nullptr;
    float *const *argOut[4] = {nullptr, nullptr, nullptr, nullptr};
    int counterSets = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    CallPointers captured;
    int graphNodes = 0;

    ~framecor_batch()
    {
        (void)hipSetDevice(device);
        if (launched) (void)hipEventSynchronize(lastLaunch);
        dropGraph();
        if (lastLaunch) (void)hipEventDestroy(lastLaunch);
        if (forked) (void)hipEventDestroy(forked);
        for (hipEvent_t e : joined) (void)hipEventDestroy(e);
        for (hipStream_t b : branch) (void)hipStreamDestroy(b);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void dropGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        captured.p.clear();
    }

    bool isInt8() const { return method == FRAMECOR_METHOD_INT8; }

    // the frames of clip c in stage s: [first, first + count)
    void stageFrames(int c, int s, int &first, int &count) const
    {
        const int G = clips[c]->groupCap;
        first = s * G;
        count = std::max(0, std::min(G, frames[c] - first));
    }

    // the kernel, block, LDS and persistent grid of a shape's plan; false if perCu exceeds what fits
    bool prepare(Shape &sh) const
    {
        if (isInt8()) {
            const Int8Kernel k = int8Kernel(sh.layout, Int8Variant(sh.variant), true);
            sh.func = k.func;
            sh.threads = k.threads;
            sh.shmem = k.shmem;
        }
        else {
            const FftKernel k = fftFastFor(sh.L) ? fftFastQueueKernel(sh.L) : fftQueueKernel(sh.L);
            sh.func = k.func;
            sh.threads = k.threads;
            sh.shmem = k.shmem;
        }
        int occ = 0;
        FC_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, sh.func, sh.threads, sh.shmem));
        occ = std::max(occ, 1);
        // spectra in a workspace: at most its blocks resident
        if (sh.work) occ = std::max(1, std::min(occ, sh.workBlocks / std::max(cus, 1)));
        sh.blocks = cus * (sh.perCu > 0 ? std::min(sh.perCu, occ) : occ);
        if (sh.work) sh.blocks = std::min(sh.blocks, sh.workBlocks);
        return sh.perCu <= occ;
    }

    int itemFrames(const Shape &sh, int c, int count) const
    {
        if (sh.fpb > 0) return std::max(1, std::min(sh.fpb, count));
        // automatic: about 2048 items per clip, as the single-clip launch
        const int nwin = clips[c]->g.nwin;
        const int nfb = std::max(1, std::min(count, (2048 + nwin - 1) / nwin));
        return (count + nfb - 1) / nfb;
    }

    // the arguments of clip c in stage s (frames, gross and outputs offset to the stage's frames)
    FrameGroup stageGroup(int c, int s, const uint8_t *fr, const int32_t *gross, float *offset, float *peak,
                          float *snr, float *cov) const
    {
        const framecor_correlator &k = *clips[c];
        int first, count;
        stageFrames(c, s, first, count);
        const size_t q0 = size_t(first) * k.g.nwin;
        FrameGroup fg{};
        fg.frames = fr + size_t(first) * k.g.pixels;
        fg.gross = gross ? gross + 2 * q0 : nullptr;
        fg.box2 = isInt8() ? k.dBox2.as<int32_t>() : nullptr;
        fg.screen = isInt8() ? k.dScreen.as<int2>() : nullptr;
        fg.sigma = isInt8() ? nullptr : k.dSigma.as<float>();
        fg.n = count;
        fg.offset = offset + 2 * q0;
        fg.peak = peak + q0;
        fg.snr = snr + q0;
        fg.cov = cov + 3 * q0;
        return fg;
    }

    // the arguments of clip c for the kernel variant of its shape (or the given variant: the layout
    // fields differ between the 16 x 16 and 32 x 32 kernels)
    Int8Args int8Args(int c, const FrameGroup &fg, int variant = -1) const
    {
        const framecor_correlator &k = *clips[c];
        Int8Args a{};
        a.g = k.g;
        a.wd = k.windowData();
        a.fg = fg;
        a.tmpl = k.dTmpl.as<uint8_t>();
        const Shape &sh = shapes[shapeOf[c]];
        const int v = variant >= 0 ? variant : sh.variant;
        fillInt8Args(a, sh.layout, int8Kernel(sh.layout, Int8Variant(v), true).variant);
        return a;
    }

    FftArgs fftArgs(int c, const FrameGroup &fg) const
    {
        const framecor_correlator &k = *clips[c];
        FftArgs a{};
        a.g = k.g;
        a.wd = k.windowData();
        a.fg = fg;
        a.spectra = k.dSpectra.as<float2>();
        a.twiddle = k.dTwiddle.as<float2>();
        // spectra beyond LDS: the shape's workspace, one per resident block
        const Shape &sh = shapes[shapeOf[c]];
        a.work = sh.work ? sh.work->as<float2>() : nullptr;
        a.workStride = fftWorkPerBlock(k.g.L) / sizeof(float2);
        a.workBlocks = sh.workBlocks;
        a.plan = k.plan;
        a.np = (k.g.L + 1) / 2;
        a.H = k.g.L / 2 + 1;
        a.rs = 2 * a.H;
        return a;
    }

    // the work queues of shape sh in stage s (at most kQueueClips clips each); fpbOverride > 0 and
    // frameCap > 0 (tuning) replace the plan's frames per item and limit the frames of each clip
    std::vector<WorkQueue> queues(const Shape &sh, int s, int fpbOverride = 0, int frameCap = 0) const
    {
        std::vector<WorkQueue> qs;
        WorkQueue q{};
        for (int c : sh.clips) {
            int first, count;
            stageFrames(c, s, first, count);
            if (frameCap > 0) count = std::min(count, frameCap);
            if (count <= 0) continue;
            if (q.nseg == kQueueClips) {
                qs.push_back(q);
                q = WorkQueue{};
            }
            Shape tmp = sh;
            if (fpbOverride > 0) tmp.fpb = fpbOverride;
            const int fpb = itemFrames(tmp, c, count), nfb = (count + fpb - 1) / fpb;
            const int64_t items = int64_t(clips[c]->g.nwin) * nfb;
            if (int64_t(q.nitems) + items > 0x7fffffff) throw std::invalid_argument("framecor: too many work items in one batch stage");
            q.clip[q.nseg] = c;
            q.first[q.nseg] = q.nitems;
            q.nfb[q.nseg] = nfb;
            q.fpb[q.nseg] = fpb;
            q.frames[q.nseg] = count;
            q.nseg++;
            q.nitems += int(items);
            q.first[q.nseg] = q.nitems;
        }
        if (q.nseg) qs.push_back(q);
        return qs;
    }

    void launchShape(const Shape &sh, const WorkQueue &q, const void *args, hipStream_t s) const
    {
        WorkQueue qa = q;
        const void *a = args;
        void *kargs[] = {&qa, &a};
        FC_HIP_CHECK(hipLaunchKernel(sh.func, dim3(sh.blocks), dim3(sh.threads), kargs, sh.shmem, s));
    }

    size_t argSize() const { return isInt8() ? sizeof(Int8Args) : sizeof(FftArgs); }

    // the device table of every clip's arguments in every stage: [stage][clip]
    void writeArgs(const uint8_t *const *fr, const int32_t *const *gross, float *const *offset, float *const *peak,
                   float *const *snr, float *const *cov)
    {
        argFr = fr;
        argGross = gross;
        argOut[0] = offset;
        argOut[1] = peak;
        argOut[2] = snr;
        argOut[3] = cov;
        const int n = int(clips.size());
        std::vector<unsigned char> host(argSize() * size_t(stages) * n);
        for (int s = 0; s < stages; s++)
            for (int c = 0; c < n; c++) {
                const FrameGroup fg = stageGroup(c, s, fr[c], gross ? gross[c] : nullptr, offset[c], peak[c], snr[c], cov[c]);
                unsigned char *dst = host.data() + argSize() * (size_t(s) * n + c);
                if (isInt8()) {
                    const Int8Args a = int8Args(c, fg);
                    std::memcpy(dst, &a, sizeof(a));
                }
                else {
                    const FftArgs a = fftArgs(c, fg);
                    std::memcpy(dst, &a, sizeof(a));
                }
            }
        dArgs.ensure(host.size());
        FC_HIP_CHECK(hipMemcpy(dArgs.ptr, host.data(), host.size(), hipMemcpyHostToDevice));
        // the box maps of each stage's launches: one job per clip of the launch with frames in the stage
        const int ns = int(shapes.size());
        boxGroups = isInt8() ? 1 : ns;
        std::vector<int> all(n);
        for (int c = 0; c < n; c++) all[c] = c;
        std::vector<BoxJob> jobs;
        jobFirst.assign(size_t(stages) * boxGroups, 0);
        jobCount.assign(size_t(stages) * boxGroups, 0);
        jobBlocks.assign(size_t(stages) * boxGroups, 0);
        for (int st = 0; st < stages; st++)
            for (int gi = 0; gi < boxGroups; gi++) {
                const int x = st * boxGroups + gi;
                jobFirst[x] = int(jobs.size());
                for (int c : isInt8() ? all : shapes[gi].clips) {
                    const framecor_correlator &k = *clips[c];
                    int first, count;
                    stageFrames(c, st, first, count);
                    if (count <= 0) continue;
                    BoxJob j{};
                    j.frames = fr[c] + size_t(first) * k.g.pixels;
                    j.box = nullptr;
                    j.box2 = isInt8() ? k.dBox2.as<int32_t>() : nullptr;
                    j.screen = isInt8() ? k.dScreen.as<int2>() : nullptr;
                    j.sigma = isInt8() ? nullptr : k.dSigma.as<float>();
                    j.n = count;
                    j.h = k.g.h;
                    j.w = k.g.w;
                    j.W = k.g.W;
                    j.first = jobBlocks[x];
                    jobBlocks[x] += boxJobBlocks(count, k.g.h, k.g.w, k.g.W);
                    jobs.push_back(j);
                    jobCount[x]++;
                }
            }
        dJobs.ensure(sizeof(BoxJob) * std::max<size_t>(jobs.size(), 1));
        if (!jobs.empty())
            FC_HIP_CHECK(hipMemcpy(dJobs.ptr, jobs.data(), sizeof(BoxJob) * jobs.size(), hipMemcpyHostToDevice));
    }

    // the box maps of launch gi of stage st (the int8 form: every clip; the FFT form: tile shape gi)
    void launchBoxGroup(int st, int gi, hipStream_t s) const
    {
        const int x = st * boxGroups + gi;
        launchBoxJobs(dJobs.as<BoxJob>() + jobFirst[x], jobCount[x], jobBlocks[x], s);
    }

    // the stage-0 arguments of the clips of shape sh for the int8 kernel variant under trial (the table
    // must not be in use: the autotuner waits for each trial)
    void writeTrialArgs(const Shape &sh, int variant)
    {
        if (!isInt8()) return;
        const int n = int(clips.size());
        for (int c : sh.clips) {
            const FrameGroup fg = stageGroup(c, 0, argFr[c], argGross ? argGross[c] : nullptr, argOut[0][c], argOut[1][c],
                                             argOut[2][c], argOut[3][c]);
            const Int8Args a = int8Args(c, fg, variant);
            FC_HIP_CHECK(hipMemcpy(static_cast<unsigned char *>(dArgs.ptr) + argSize() * size_t(c), &a, sizeof(a),
                                   hipMemcpyHostToDevice));
        }
        (void)n;
    }

    const void *stageArgs(int s) const
    {
        return static_cast<const unsigned char *>(dArgs.ptr) + argSize() * size_t(s) * clips.size();
    }

    // the queue counters of one call (before the capture)
    void prepareCall()
    {
        int sets = 0;
        for (int st = 0; st < stages; st++)
            for (const Shape &sh : shapes) sets += int(queues(sh, st).size());
        dCounters.ensure(sizeof(int) * kXcds * std::max(sets, 1));
        counterSets = sets;
        for (Shape &sh : shapes) prepare(sh);
    }

    // the FFT form: the branch of each shape, largest first (transform work of a call) to the least
    // loaded branch; the streams and events of the branches in use (created once)
    void assignBranches()
    {
        if (isInt8()) return;
        const int ns = int(shapes.size());
        std::vector<double> work(ns, 0.0);
        for (int si = 0; si < ns; si++)
            for (int c : shapes[si].clips) {
                const Geometry &g = clips[c]->g;
                work[si] += double(g.nwin) * frames[c] * double(g.L) * g.L;
            }
        std::vector<int> order(ns);
        for (int si = 0; si < ns; si++) order[si] = si;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
        const int nb = std::min(ns, kBranches);
        std::vector<double> load(nb, 0.0);
        branchOf.assign(ns, 0);
        for (int si : order) {
            const int b = int(std::min_element(load.begin(), load.end()) - load.begin());
            branchOf[si] = b;
            load[b] += work[si];
        }
        if (!forked) FC_HIP_CHECK(hipEventCreateWithFlags(&forked, hipEventDisableTiming));
        while (int(branch.size()) < nb) {
            hipStream_t b = nullptr;
            FC_HIP_CHECK(hipStreamCreateWithFlags(&b, hipStreamNonBlocking));
            branch.push_back(b);
            hipEvent_t e = nullptr;
            FC_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            joined.push_back(e);
        }
    }

    // the launch sequence of one call (captured into the graph): the int8 form on s alone; the FFT form
    // every shape's stages on its branch, forked from s after the counters are zeroed and joined to s at
    // the end
    void enqueueCall(hipStream_t s) const
    {
        const int sets = counterSets;
        int *counters = dCounters.as<int>();
        hipLaunchKernelGGL(zeroInts, dim3(std::max(1, (kXcds * sets + 255) / 256)), dim3(256), 0, s, counters, kXcds * sets);
        FC_HIP_CHECK(hipGetLastError());
        if (isInt8()) {
            int set = 0;
            for (int st = 0; st < stages; st++) {
                launchBoxGroup(st, 0, s);
                for (const Shape &sh : shapes)
                    for (WorkQueue q : queues(sh, st)) {
                        q.counters = counters + kXcds * set++;
                        launchShape(sh, q, stageArgs(st), s);
                    }
            }
            return;
        }
        const int ns = int(shapes.size()), nb = std::min(ns, kBranches);
        FC_HIP_CHECK(hipEventRecord(forked, s));
        for (int b = 0; b < nb; b++) FC_HIP_CHECK(hipStreamWaitEvent(branch[b], forked, 0));
        int set = 0;
        for (int si = 0; si < ns; si++) {
            const hipStream_t bs = branch[branchOf[si]];
            for (int st = 0; st < stages; st++) {
                launchBoxGroup(st, si, bs);
                for (WorkQueue q : queues(shapes[si], st)) {
                    q.counters = counters + kXcds * set++;
                    launchShape(shapes[si], q, stageArgs(st), bs);
                }
            }
        }
        for (int b = 0; b < nb; b++) {
            FC_HIP_CHECK(hipEventRecord(joined[b], branch[b]));
            FC_HIP_CHECK(hipStreamWaitEvent(s, joined[b], 0));
        }
    }

    // times the candidates of every shape without a plan on samples of stage 0 and keeps the fastest
    void autotune()
    {
        hipEvent_t e0, e1;
        FC_HIP_CHECK(hipEventCreate(&e0));
        FC_HIP_CHECK(hipEventCreate(&e1));
        int *counters = nullptr;
        DeviceBuffer dTune;
        dTune.ensure(sizeof(int) * kXcds * 64);
        counters = dTune.as<int>();
        try {
            // the box sums of stage 0 (the samples read them)
            for (int gi = 0; gi < boxGroups; gi++) launchBoxGroup(0, gi, stream);
            for (Shape &sh : shapes) {
                if (sh.tuned) continue;
                // a sample of about 64 items per CU, and at least as many frames as the largest frames per
                // item (else every candidate of fpb above the sample's frames runs the same work, and
                // noise picks among them); at most the stage's frames of each clip
                int64_t perFrame = 0;
                for (int c : sh.clips) perFrame += clips[c]->g.nwin;
                const int cap = int(std::max<int64_t>(32, std::min<int64_t>(64, (int64_t(cus) * 64 + perFrame - 1) / perFrame)));
                std::vector<int> variants;
                if (isInt8()) {
                    if (sh.layout.quadOk) variants.push_back(int(Int8Variant::Quad));
                    if (sh.layout.wideOk) variants.push_back(int(Int8Variant::Wide));
                    if (sh.layout.edge) variants.push_back(int(Int8Variant::Simple));
                    variants.push_back(int(Int8Variant::Padded));
                    if (sh.layout.pipedOk) variants.push_back(int(Int8Variant::Piped));
                    if (sh.layout.soloOk) variants.push_back(int(Int8Variant::Solo));
                }
                else {
                    variants.push_back(int(Int8Variant::Auto));
                }
                float best = 1e30f;
                Shape keep = sh;
                int lastTrial = -1;
                // resident blocks per CU: one-wave blocks 8-16, blocks of four or more waves 1-4
                static const std::vector<int> fewBlocks{1, 2, 3, 4}, manyBlocks{8, 12, 16};
                for (int v : variants)
                    for (int fpb : {1, 2, 4, 8, 16, 32})
                        for (int perCu : v == int(Int8Variant::Solo) ? manyBlocks : fewBlocks) {
                            Shape cand = sh;
                            cand.variant = v;
                            if (cand.variant != lastTrial) {
                                // the previous trial's kernels are over (each trial waits for its events)
                                writeTrialArgs(cand, cand.variant);
                                lastTrial = cand.variant;
                            }
                            cand.fpb = fpb;
                            cand.perCu = perCu;
                            // grids beyond the occupancy are the launch of a smaller perCu
                            if (!prepare(cand)) continue;
                            const std::vector<WorkQueue> qs = queues(cand, 0, fpb, cap);
                            if (qs.empty()) break;
                            float t = 1e30f;
                            for (int rep = 0; rep < 2; rep++) {
                                int nq = int(qs.size());
                                hipLaunchKernelGGL(zeroInts, dim3(1), dim3(256), 0, stream, counters, kXcds * std::min(nq, 64));
                                FC_HIP_CHECK(hipEventRecord(e0, stream));
                                for (int i = 0; i < nq && i < 64; i++) {
                                    WorkQueue q = qs[i];
                                    q.counters = counters + kXcds * i;
                                    launchShape(cand, q, stageArgs(0), stream);
                                }
                                FC_HIP_CHECK(hipEventRecord(e1, stream));
                                FC_HIP_CHECK(hipEventSynchronize(e1));
                                float ms = 0.0f;
                                FC_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                                t = std::min(t, ms);
                            }
                            if (t < best) {
                                best = t;
                                keep = cand;
                            }
                        }
                sh.variant = keep.variant;
                sh.fpb = keep.fpb;
                sh.perCu = keep.perCu;
                sh.tuned = true;
            }
        }
        catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        FC_HIP_CHECK(hipEventDestroy(e0));
        FC_HIP_CHECK(hipEventDestroy(e1));
    }

    void correlate(const uint8_t *const *fr, const int32_t *const *gross, float *const *offset, float *const *peak,
                   float *const *snr, float *const *cov, hipStream_t user)
    {
        const int n = int(clips.size());
        for (int c = 0; c < n; c++) {
            if (!clips[c]->ready) throw std::runtime_error("framecor: set_reference first (clip " + std::to_string(c) + ")");
            if (frames[c] > 0 && (!fr[c] || !offset[c] || !peak[c] || !snr[c] || !cov[c]))
                invalid("null argument (clip " + std::to_string(c) + ")");
        }
        FC_HIP_CHECK(hipSetDevice(device));
        CallPointers now;
        for (int c = 0; c < n; c++)
            for (const void *p : {static_cast<const void *>(fr[c]), static_cast<const void *>(gross ? gross[c] : nullptr),
                                  static_cast<const void *>(offset[c]), static_cast<const void *>(peak[c]),
                                  static_cast<const void *>(snr[c]), static_cast<const void *>(cov[c])})
                now.p.push_back(p);
        bool needTune = false;
        for (const Shape &sh : shapes) needTune |= tune && !sh.tuned;
        if (!exec || !(now == captured) || needTune) {
            // the previous launch reads the argument table: let it finish before the table changes
            if (launched) FC_HIP_CHECK(hipEventSynchronize(lastLaunch));
            dropGraph();
            writeArgs(fr, gross, offset, peak, snr, cov);
            // the frames come from the caller's stream
            FC_HIP_CHECK(hipEventRecord(lastLaunch, user));
            FC_HIP_CHECK(hipStreamWaitEvent(stream, lastLaunch, 0));
            if (needTune) {
                autotune();
                // the tuned variants' layout fields
                writeArgs(fr, gross, offset, peak, snr, cov);
            }
            prepareCall();
            assignBranches();
            FC_HIP_CHECK(hipStreamSynchronize(stream));
            FC_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
            try {
                enqueueCall(stream);
            }
            catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            FC_HIP_CHECK(hipStreamEndCapture(stream, &graph));
            FC_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            size_t nodes = 0;
            FC_HIP_CHECK(hipGraphGetNodes(graph, nullptr, &nodes));
            graphNodes = int(nodes);
            captured = now;
        }
        FC_HIP_CHECK(hipGraphLaunch(exec, user));
        FC_HIP_CHECK(hipEventRecord(lastLaunch, user));
        launched = true;
    }
};

extern "C" {

int framecor_batch_create(const framecor_clip *clips, int nclips, int method, int device, framecor_batch **out)
{
    if (!out) return fail(FRAMECOR_EINVAL, "framecor: no output handle");
    *out = nullptr;
    return guarded([&] {
        if (nclips <= 0 || !clips) invalid("a batch needs at least one clip");
        if (method != FRAMECOR_METHOD_FFT && method != FRAMECOR_METHOD_INT8) invalid("the method must be fft or int8");
        // every clip's parameters first, before any GPU call
        for (int c = 0; c < nclips; c++) {
            const framecor_clip &k = clips[c];
            try {
                checkAll(k.height, k.width, k.down, k.ndown, k.across, k.nacross, k.window, k.search, method, 1,
                         k.stat_size);
            }
            catch (const std::invalid_argument &e) {
                throw std::invalid_argument(std::string(e.what()) + " (clip " + std::to_string(c) + ")");
            }
            if (k.frames < 0) invalid("the number of frames must not be negative (clip " + std::to_string(c) + ")");
        }
        auto b = std::make_unique<framecor_batch>();
        b->method = method;
        b->device = device;
        for (int c = 0; c < nclips; c++) {
            const framecor_clip &k = clips[c];
            auto cor = std::make_unique<framecor_correlator>();
            cor->init(k.height, k.width, k.down, k.ndown, k.across, k.nacross, k.window, k.search, method, 1,
                      k.stat_size, device);
            b->clips.push_back(std::move(cor));
            b->frames.push_back(k.frames);
        }
        FC_HIP_CHECK(hipSetDevice(device));
        // the box maps of every clip's frames in one stage where they fit in a quarter of the free device
        // memory (at most 64 GiB); else each clip's frames in groups of its own limit (2 GiB of maps)
        size_t mapBytes = 0, freeBytes = 0, totalBytes = 0;
        for (int c = 0; c < nclips; c++) mapBytes += size_t(clips[c].frames) * b->clips[c]->bytesPerFrame();
        FC_HIP_CHECK(hipMemGetInfo(&freeBytes, &totalBytes));
        const bool oneStage = mapBytes <= std::min(freeBytes / 4, size_t(64) << 30);
        for (int c = 0; c < nclips; c++) {
            framecor_correlator &k = *b->clips[c];
            if (oneStage) k.maxGroup = std::max(k.maxGroup, clips[c].frames);
            if (clips[c].frames > 0) k.ensureGroup(clips[c].frames);
        }
        FC_HIP_CHECK(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device));
        FC_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        FC_HIP_CHECK(hipEventCreateWithFlags(&b->lastLaunch, hipEventDisableTiming));
        // tile shapes
        b->shapeOf.assign(nclips, -1);
        for (int c = 0; c < nclips; c++) {
            const Geometry &g = b->clips[c]->g;
            int s = 0;
            for (; s < int(b->shapes.size()); s++) {
                const Shape &sh = b->shapes[s];
                // int8: one kernel and one LDS layout per (window, search range, SNR box side)
                const int snr = std::min(g.stat, g.M);
                if (b->isInt8() ? (sh.W == g.W && sh.R == g.R && sh.snr == snr) : sh.L == g.L) break;
            }
            if (s == int(b->shapes.size())) {
                Shape sh;
                sh.W = g.W;
                sh.R = g.R;
                sh.L = g.L;
                sh.snr = std::min(g.stat, g.M);
                if (b->isInt8()) sh.layout = b->clips[c]->i8;
                else if (!fftInLds(g.L)) {
                    sh.workBlocks = fftWorkBlocks(g.L, b->cus);
                    sh.work = std::make_shared<DeviceBuffer>();
                    sh.work->ensure(fftWorkPerBlock(g.L) * size_t(sh.workBlocks));
                }
                b->shapes.push_back(sh);
            }
            b->shapes[s].clips.push_back(c);
            b->shapeOf[c] = s;
        }
        // stages: each clip's frames in groups of what its buffers hold
        for (int c = 0; c < nclips; c++) {
            const int G = std::max(1, b->clips[c]->groupCap);
            b->stages = std::max(b->stages, (b->frames[c] + G - 1) / G);
        }
        *out = b.release();
    });
}

void framecor_batch_destroy(framecor_batch *b) { delete b; }

int framecor_batch_set_reference(framecor_batch *b, int clip, const uint8_t *frame)
{
    if (!b || !frame) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    if (clip < 0 || clip >= int(b->clips.size())) return fail(FRAMECOR_EINVAL, "framecor: no clip " + std::to_string(clip));
    return guarded([&] {
        if (b->launched) FC_HIP_CHECK(hipEventSynchronize(b->lastLaunch));
        b->clips[clip]->setReference(frame);
    });
}

int framecor_batch_set_tuning(framecor_batch *b, int enable)
{
    if (!b) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    b->tune = enable != 0;
    return FRAMECOR_OK;
}

int framecor_batch_plan(const framecor_batch *b, int clip, int *plan)
{
    if (!b || !plan) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    if (clip < 0 || clip >= int(b->clips.size())) return fail(FRAMECOR_EINVAL, "framecor: no clip " + std::to_string(clip));
    const Shape &sh = b->shapes[b->shapeOf[clip]];
    int variant = sh.variant;
    if (b->isInt8()) variant = int(int8Kernel(sh.layout, Int8Variant(sh.variant), true).variant);
    plan[0] = variant;
    plan[1] = sh.fpb;
    plan[2] = sh.perCu;
    plan[3] = sh.tuned ? 1 : 0;
    return FRAMECOR_OK;
}

int framecor_batch_set_plan(framecor_batch *b, int clip, const int *plan)
{
    if (!b || !plan) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    if (clip < 0 || clip >= int(b->clips.size())) return fail(FRAMECOR_EINVAL, "framecor: no clip " + std::to_string(clip));
    if (plan[0] < 0 || plan[0] > 6 || plan[1] < 0 || plan[2] < 0 || plan[2] > 64)
        return fail(FRAMECOR_EINVAL, "framecor: a plan is (variant 0..6, frames per item >= 0, blocks per CU 0..64)");
    return guarded([&] {
        if (b->launched) FC_HIP_CHECK(hipEventSynchronize(b->lastLaunch));
        Shape &sh = b->shapes[b->shapeOf[clip]];
        sh.variant = plan[0];
        sh.fpb = plan[1];
        sh.perCu = plan[2];
        sh.tuned = true;
        b->dropGraph();
    });
}

int framecor_batch_set_stage_frames(framecor_batch *b, int frames)
{
    if (!b) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    if (frames < 1) return fail(FRAMECOR_EINVAL, "framecor: a stage needs at least one frame per clip");
    return guarded([&] {
        if (b->launched) FC_HIP_CHECK(hipEventSynchronize(b->lastLaunch));
        b->stages = 0;
        for (size_t c = 0; c < b->clips.size(); c++) {
            framecor_correlator &k = *b->clips[c];
            k.groupCap = std::min(k.groupCap, frames);
            const int G = std::max(1, k.groupCap);
            b->stages = std::max(b->stages, (b->frames[c] + G - 1) / G);
        }
        b->dropGraph();
    });
}

int framecor_batch_info(const framecor_batch *b, int *info)
{
    if (!b || !info) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    info[0] = int(b->shapes.size());
    info[1] = b->stages;
    info[2] = b->graphNodes;
    info[3] = b->counterSets;
    return FRAMECOR_OK;
}

int framecor_batch_correlate_device(framecor_batch *b, const uint8_t *const *frames, const int32_t *const *gross,
                                    float *const *offset, float *const *peak, float *const *snr, float *const *cov,
                                    void *stream)
{
    if (!b || !frames || !offset || !peak || !snr || !cov) return fail(FRAMECOR_EINVAL, "framecor: null argument");
    return guarded([&] { b->correlate(frames, gross, offset, peak, snr, cov, static_cast<hipStream_t>(stream)); });
}

} // extern "C"
