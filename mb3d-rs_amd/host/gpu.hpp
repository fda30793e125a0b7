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

}  // namespace mb3d::gpu
