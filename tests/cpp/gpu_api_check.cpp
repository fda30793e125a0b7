// A C++ caller of the backend's `gpu` interface (mb3d-rs_amd/host/gpu.hpp),
// driven by tests/test_host_cpp.py.
//
//   gpu_api_check cpu
//       the parts that need no device: record packing, the switch, the
//       status texts, unsupported scenes
//   gpu_api_check gpu PARAMS POST OUTDIR
//       marches the picture of PARAMS (raw mb3d_hip_params bytes; POST a
//       raw mb3d_hip_post_job or "-") three ways -- gpu::march with its
//       sink, gpu::march_bands over two bands, mb3d_hip_march directly --
//       and writes each G-buffer (5 words per pixel) to OUTDIR, plus the
//       post passes of gpu::post_on_card on the plain march's G-buffer
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <thread>
#include <vector>

#include "../../mb3d-rs_amd/host/gpu.hpp"

namespace g = mb3d::gpu;

static int fail(const char *what) {
    std::fprintf(stderr, "FAIL: %s\n", what);
    return 1;
}

static std::vector<char> read_file(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    return std::vector<char>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

static void write_words(const std::string &path, const std::vector<g::SiLight> &gbuf) {
    std::vector<uint32_t> w(gbuf.size() * 5);
    for (size_t i = 0; i < gbuf.size(); i++) g::pack(gbuf[i], &w[i * 5]);
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(w.data()), (std::streamsize)(w.size() * 4));
}

static int cpu_checks() {
    // the card record round trip, every field at a distinct value
    g::SiLight s;
    s.normal[0] = -12345;
    s.normal[1] = 32767;
    s.normal[2] = -32767;
    s.zpos_fine = 0x12345678u;
    s.shadow = 0xFC3Fu;
    s.amb_shadow = 16383;
    s.si_gradient = 0x8001u;
    s.otrap = 4096;
    uint32_t w[5];
    g::pack(s, w);
    const g::SiLight t = g::unpack(w);
    if (std::memcmp(t.normal, s.normal, sizeof s.normal) || t.zpos_fine != s.zpos_fine || t.shadow != s.shadow ||
        t.amb_shadow != s.amb_shadow || t.si_gradient != s.si_gradient || t.otrap != s.otrap)
        return fail("SiLight pack/unpack");
    if (g::SiLight().zpos_fine >> 16 != 32768u) return fail("the default record is not background");

    // a scene the backend lacks: a reason, and march says CPU
    mb3d_hip_scene sc;
    mb3d_hip_scene_default(&sc);
    sc.width = 32;
    sc.height = 24;
    mb3d_hip_params p;
    if (mb3d_hip_params_from_scene(&sc, &p) != MB3D_HIP_OK) return fail("params_from_scene");
    if (g::unsupported(p)) return fail("the default scene is supported");
    mb3d_hip_params bad = p;
    bad.slot_kind[0] = 99;
    const auto why = g::unsupported(bad);
    if (!why || why->empty()) return fail("an unknown formula has no reason");

    // switched off: no device is touched, the status says why
    g::set_enabled(false);
    if (g::enabled()) return fail("set_enabled(false)");
    bool sunk = false;
    const auto r = g::march(p, nullptr, nullptr, nullptr, [&](size_t, const g::SiLight *, size_t) { sunk = true; });
    if (r || sunk) return fail("march ran while switched off");
    if (g::last_status().rfind("CPU: ", 0) != 0 || g::thread_status() != g::last_status())
        return fail("the status of a switched-off march");
    const g::Status st = g::march_bands(p, nullptr, [](size_t) { return std::nullopt; }, nullptr, nullptr);
    if (st.ok) return fail("march_bands ran while switched off");
    g::set_enabled(true);
    if (g::rows_overflowed()) return fail("rows_overflowed without a march");
    if (g::aborted()) return fail("aborted before abort_all");
    std::printf("ok\n");
    return 0;
}

static int gpu_checks(const std::string &params_path, const std::string &post_path, const std::string &outdir) {
    const std::vector<char> pb = read_file(params_path);
    if (pb.size() != sizeof(mb3d_hip_params)) return fail("params size");
    mb3d_hip_params p;
    std::memcpy(&p, pb.data(), sizeof p);
    mb3d_hip_post_job job;
    const mb3d_hip_post_job *post = nullptr;
    if (post_path != "-") {
        const std::vector<char> jb = read_file(post_path);
        if (jb.size() != sizeof job) return fail("post job size");
        std::memcpy(&job, jb.data(), sizeof job);
        post = &job;
    }
    const size_t w = (size_t)p.width, h = (size_t)p.height;

    // gpu::march: rows in order to the sink, progress to the end
    std::vector<g::SiLight> sunk(w * h);
    size_t next_row = 0, last_done = 0;
    bool in_order = true;
    const auto whole = g::march(
        p, post, [&](size_t d, size_t total) { last_done = d == total ? d : last_done; }, [] { return false; },
        [&](size_t y, const g::SiLight *px, size_t n) {
            in_order &= y == next_row && n == w;
            next_row = y + 1;
            std::copy(px, px + n, sunk.begin() + y * w);
        });
    if (!whole) {
        std::fprintf(stderr, "status: %s\n", g::last_status().c_str());
        return fail("gpu::march");
    }
    if (!in_order || next_row != h || last_done != h) return fail("the sink's rows or the progress");
    if (g::last_status().rfind("GPU", 0) != 0) return fail("the status of a card march");
    for (size_t i = 0; i < w * h; i++)
        if (std::memcmp(&sunk[i], &(*whole)[i], sizeof(g::SiLight))) return fail("sink rows differ from the result");
    write_words(outdir + "/march.bin", *whole);

    // gpu::march_bands: the lower half first, then the upper
    std::vector<g::SiLight> banded(w * h);
    const size_t half = h / 2;
    const g::Status st = g::march_bands(
        p, post,
        [&](size_t i) -> std::optional<std::pair<size_t, size_t>> {
            if (i == 0) return std::make_pair(half, h - half);
            if (i == 1) return std::make_pair((size_t)0, half);
            return std::nullopt;
        },
        [] { return false; },
        [&](size_t y, const g::SiLight *px, size_t n) { std::copy(px, px + n, banded.begin() + y * w); });
    if (!st.ok) return fail(st.error.c_str());
    write_words(outdir + "/bands.bin", banded);

    // the C ABI directly
    std::vector<uint32_t> words(w * h * 5);
    mb3d_hip_out out;
    std::memset(&out, 0, sizeof out);
    out.gbuf = words.data();
    if (mb3d_hip_march(&p, post, 0, (int32_t)h, &out, nullptr, nullptr, nullptr, nullptr) != MB3D_HIP_OK)
        return fail("mb3d_hip_march");
    {
        std::ofstream f(outdir + "/direct.bin", std::ios::binary);
        f.write(reinterpret_cast<const char *>(words.data()), (std::streamsize)(words.size() * 4));
    }

    // gpu::post_on_card on the G-buffer of a march without post passes
    if (post) {
        auto plain = g::march(p, nullptr, nullptr, nullptr, nullptr);
        if (!plain) return fail("plain march");
        const g::Status ps = g::post_on_card(p, *plain, *post, nullptr, [] { return false; });
        if (!ps.ok) return fail(ps.error.c_str());
        write_words(outdir + "/post.bin", *plain);
    }
    std::printf("ok\n");
    return 0;
}

// gpu::march_mixed: the CPU marcher is the picture the test computed
// beforehand (CPU_GBUF, 5 words a pixel), one row per call after `delay_us`;
// rows y % 7 == 3 are a checkpoint's (a marker record).  Writes the result,
// each row's sink count, and the rows the card and the CPU delivered.
static int mixed_checks(const std::string &params_path, const std::string &cpu_path, const std::string &outdir,
                        int threads, int delay_us, int stop_after) {
    const auto pb = read_file(params_path);
    mb3d_hip_params p;
    if (pb.size() != sizeof p) return fail("params size");
    std::memcpy(&p, pb.data(), sizeof p);
    const size_t w = (size_t)p.width, h = (size_t)p.height;
    const auto cb = read_file(cpu_path);
    if (cb.size() != w * h * 20) return fail("the CPU picture's size");
    std::vector<uint32_t> cpu_words(w * h * 5);
    std::memcpy(cpu_words.data(), cb.data(), cb.size());
    g::SiLight marker;
    marker.zpos_fine = 0x5A5A5A5Au;
    marker.si_gradient = 0x1234;
    std::vector<g::SiLight> have(w * h, marker);
    std::vector<int> sinks(h, 0);
    std::atomic<size_t> cpu_calls{0};
    g::CpuRow cpu_row = [&](size_t y, g::SiLight *row) {
        if (delay_us > 0) std::this_thread::sleep_for(std::chrono::microseconds(delay_us));
        for (size_t x = 0; x < w; x++) row[x] = g::unpack(&cpu_words[(y * w + x) * 5]);
        cpu_calls++;
    };
    std::atomic<size_t> delivered{0};
    const auto res = g::march_mixed(
        p, nullptr, (size_t)threads, cpu_row,
        [&](size_t y, const g::SiLight *, size_t n) {
            if (n != w) std::abort();
            sinks[y]++;
            delivered++;
        },
        [&] { return stop_after > 0 && delivered.load() >= (size_t)stop_after; }, nullptr,
        [](size_t y) { return y % 7 == 3; }, &have, 8);
    write_words(outdir + "/mixed.bin", res.gbuf);
    {
        std::ofstream f(outdir + "/sinks.bin", std::ios::binary);
        f.write(reinterpret_cast<const char *>(sinks.data()), (std::streamsize)(sinks.size() * sizeof(int)));
    }
    std::printf("card_rows %zu cpu_rows %zu cpu_calls %zu status %s %s\n", res.card_rows, res.cpu_rows,
                cpu_calls.load(), res.status.ok ? "ok" : res.status.error.c_str(), g::last_status().c_str());
    std::printf("ok\n");
    return 0;
}

// gpu::mc_batch_shared: the CPU marcher is the oracle's results the test
// computed beforehand (CPU_RES, one mb3d_hip_mc_res a task), copied after
// `delay_us` per chunk.  Writes the results and each task's side.
static int mc_share_checks(const std::string &params_path, const std::string &tasks_path, const std::string &cpu_path,
                           const std::string &outdir, int threads, int delay_us, int chunk) {
    const auto pb = read_file(params_path);
    mb3d_hip_params p;
    if (pb.size() != sizeof p) return fail("params size");
    std::memcpy(&p, pb.data(), sizeof p);
    const auto tb = read_file(tasks_path);
    const auto cb = read_file(cpu_path);
    const size_t n = tb.size() / sizeof(mb3d_hip_mc_task);
    if (tb.size() != n * sizeof(mb3d_hip_mc_task) || cb.size() != n * sizeof(mb3d_hip_mc_res))
        return fail("tasks or results size");
    std::vector<mb3d_hip_mc_task> tasks(n);
    std::memcpy(tasks.data(), tb.data(), tb.size());
    const auto *want = reinterpret_cast<const mb3d_hip_mc_res *>(cb.data());
    std::atomic<size_t> calls{0};
    g::CpuTasks cpu = [&](const mb3d_hip_mc_task *t, size_t m, mb3d_hip_mc_res *r) {
        if (delay_us > 0) std::this_thread::sleep_for(std::chrono::microseconds(delay_us));
        const size_t i0 = (size_t)(t - tasks.data());
        std::memcpy(r, want + i0, m * sizeof *r);
        calls++;
    };
    const auto res = g::mc_batch_shared(p, tasks, false, (size_t)threads, cpu, (size_t)chunk, 64);
    {
        std::ofstream f(outdir + "/mc_res.bin", std::ios::binary);
        f.write(reinterpret_cast<const char *>(res.res.data()), (std::streamsize)(n * sizeof(mb3d_hip_mc_res)));
        std::ofstream s(outdir + "/mc_side.bin", std::ios::binary);
        s.write(reinterpret_cast<const char *>(res.side.data()), (std::streamsize)n);
    }
    std::printf("card %zu cpu %zu calls %zu status %s\n", res.card, res.cpu, calls.load(),
                res.status.ok ? "ok" : res.status.error.c_str());
    std::printf("ok\n");
    return 0;
}

int main(int argc, char **argv) {
    if (argc >= 2 && std::string(argv[1]) == "cpu") return cpu_checks();
    if (argc >= 5 && std::string(argv[1]) == "gpu") return gpu_checks(argv[2], argv[3], argv[4]);
    if (argc >= 9 && std::string(argv[1]) == "mcshare")
        return mc_share_checks(argv[2], argv[3], argv[4], argv[5], std::atoi(argv[6]), std::atoi(argv[7]),
                               std::atoi(argv[8]));
    if (argc >= 8 && std::string(argv[1]) == "mixed")
        return mixed_checks(argv[2], argv[3], argv[4], std::atoi(argv[5]), std::atoi(argv[6]), std::atoi(argv[7]));
    std::fprintf(stderr, "usage: gpu_api_check cpu | gpu PARAMS POST|- OUTDIR | mixed PARAMS CPU_GBUF OUTDIR THREADS "
                         "DELAY_US STOP_AFTER\n");
    return 2;
}
