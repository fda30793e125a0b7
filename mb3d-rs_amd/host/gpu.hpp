// gpu.hpp -- the reference's `crate::gpu` interface for this backend, in
// C++ over the C ABI (include/mb3d_hip.h): the same names, argument meaning
// and error behaviour as src/gpu.rs, so that a C++ renderer (or the tests)
// drive the MI355X marcher the way src/render.rs drives the wgpu one.
//
//   march        gpu::march          (src/gpu.rs:2296)  None -> calculate on the CPU
//   march_bands  gpu::march_bands    (src/gpu.rs:1586)  rows of the bands handed out
//   post_on_card gpu::post_on_card   (src/gpu.rs:1649)
//   unsupported  gpu::unsupported    (src/gpu.rs:445)
//   enabled / set_enabled / last_status / set_status / thread_status /
//   rows_overflowed / clear_rows_overflowed / abort_all / aborted
//                                    (src/gpu.rs:282-443, 2278-2286)
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <optional>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mb3d_hip.h"

namespace mb3d::gpu {

// The G-buffer record (`SiLight`, src/gbuffer.rs:4-20).
struct SiLight {
    int16_t normal[3] = {0, 0, 0};
    uint32_t zpos_fine = 32768u << 16;  // background
    uint16_t shadow = 0;
    uint16_t amb_shadow = 0;
    uint16_t si_gradient = 0;
    uint16_t otrap = 0;
};

// The card's 20-byte record of one pixel (src/gpu.rs:3452-3463) and back.
SiLight unpack(const uint32_t w[5]);
void pack(const SiLight &s, uint32_t w[5]);

using Progress = std::function<void(size_t done, size_t total)>;
using Cancel = std::function<bool()>;
using Sink = std::function<void(size_t row, const SiLight *pixels, size_t n)>;
// the next band (row0, rows) for band index i, or none when all are handed out
using Bands = std::function<std::optional<std::pair<size_t, size_t>>(size_t i)>;

// Ok or an error text (Result<(), String>).
struct Status {
    bool ok = true;
    std::string error;
    static Status fail(std::string e) { return Status{false, std::move(e)}; }
};

// Whether the card is used (the programs' switch; on unless MB3D_GPU=0).
bool enabled();
void set_enabled(bool on);
// How the last calculation ran: "GPU (name)" or "CPU: reason".
std::string last_status();
// last_status as the calling thread set it.
std::string thread_status();
void set_status(const std::string &s);
// The last march's rows were given up for single-precision overflow (the
// caller redoes them on the CPU).
bool rows_overflowed();
void clear_rows_overflowed();
// Stops all card work for good; aborted() is then true.
void abort_all();
bool aborted();

// Why the backend cannot march `p` (none when it can).
std::optional<std::string> unsupported(const mb3d_hip_params &p);

// The G-buffer of the picture with the post passes of `post` (may be
// null); rows go to `sink` in order as bands finish.  None when the scene or
// the computer cannot use the card, when single precision overflowed, or
// when cancelled; the reason is in last_status().
std::optional<std::vector<SiLight>> march(const mb3d_hip_params &p, const mb3d_hip_post_job *post,
                                          const Progress &progress, const Cancel &cancel, const Sink &sink,
                                          const mb3d_hip_options *options = nullptr);

// The bands that `bands` hands out, with the post passes on each; their
// rows go to `sink`.  An error after rows went to `sink` (overflow) means
// they are not usable.
Status march_bands(const mb3d_hip_params &p, const mb3d_hip_post_job *post, const Bands &bands,
                   const Cancel &cancel, const Sink &sink, const mb3d_hip_options *options = nullptr);

// The post passes of `post` on a G-buffer marched elsewhere, band by band.
Status post_on_card(const mb3d_hip_params &p, std::vector<SiLight> &gbuf, const mb3d_hip_post_job &post,
                    const Progress &progress, const Cancel &cancel);

// ---- the march shared between the card and the CPU's threads
// (render::march_mixed and its Board, src/render.rs:1380-1470, 1500-1700;
// the renderer's Auto mode) ----

// The caller's CPU marcher: row y of the picture (width pixels) into `row`.
using CpuRow = std::function<void(size_t y, SiLight *row)>;
// A checkpoint holds row y already (its pixels are in `have`).
using Done = std::function<bool(size_t y)>;

struct Mixed {
    std::vector<SiLight> gbuf;
    size_t card_rows = 0;  // rows the card delivered (0: it never joined, or its rows were redone)
    size_t cpu_rows = 0;   // rows the CPU's threads delivered
    Status status;         // fail("cancelled") when stopped
};

// The picture's rows marched by `threads` CPU threads (cpu_row), from the
// bottom, and by the card, in bands of `band_rows` from the top (with the
// post passes of `post` on its bands), until they meet.  Rows that `done`
// names are the checkpoint's, taken from `have` and not marched.  When no
// row is free, an idle worker backs up rows the other side took and has not
// delivered; the first result of a row counts.  Every row reaches `sink`
// as it is delivered, once -- but for a card whose bands overflow single
// precision: its rows are then redone by the CPU marcher and reach `sink`
// again.  A card that cannot start leaves all rows to the CPU threads.
Mixed march_mixed(const mb3d_hip_params &p, const mb3d_hip_post_job *post, size_t threads, const CpuRow &cpu_row,
                  const Sink &sink, const Cancel &cancel, const Progress &progress, const Done &done = {},
                  const std::vector<SiLight> *have = nullptr, size_t band_rows = 16,
                  const mb3d_hip_options *options = nullptr);

// ---- the Monte Carlo renderer's batches shared with the CPU's threads
// (gpu::mc_batch with RaySharing, src/gpu.rs:3013-3190: the card starts the
// tasks from the front, the CPU takes them from the end) ----

// The caller's CPU marcher of n Monte Carlo tasks (McMarch::run) into `res`.
using CpuTasks = std::function<void(const mb3d_hip_mc_task *tasks, size_t n, mb3d_hip_mc_res *res)>;

struct McShared {
    std::vector<mb3d_hip_mc_res> res;
    std::vector<uint8_t> side;  // per task: 1 the card marched it, 2 the CPU, 0 nobody (cancelled)
    size_t card = 0, cpu = 0;   // tasks each side marched
    Status status;              // fail("cancelled") when stopped, or the card's error
};

// The batch marched by the card in chunks of `chunk` tasks from the front
// and by `threads` CPU threads (cpu) in chunks of `cpu_chunk` from the end,
// until they meet: every task once, by one side.  A card that fails or
// cannot start leaves the rest to the CPU threads (status then names its
// error, and every task is still marched).
McShared mc_batch_shared(const mb3d_hip_params &p, const std::vector<mb3d_hip_mc_task> &tasks, bool secant,
                         size_t threads, const CpuTasks &cpu, size_t chunk = 65536, size_t cpu_chunk = 256,
                         const Cancel &cancel = {});

}  // namespace mb3d::gpu
