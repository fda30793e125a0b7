// gpu.cpp -- the reference's `crate::gpu` interface over the C ABI (see
// gpu.hpp).  Bands of rows go through mb3d_hip_march one after another, so
// that rows reach the sink, and progress and cancel are asked, between
// bands (as the wgpu backend's slices, src/gpu.rs:3215-3486).
#include "gpu.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>

namespace mb3d::gpu {

namespace {

constexpr size_t BAND_ROWS = 128;  // rows per call between sink, progress and cancel

std::atomic<bool> &enabled_flag() {
    static std::atomic<bool> on([] {
        const char *e = std::getenv("MB3D_GPU");
        return !(e && std::strcmp(e, "0") == 0);
    }());
    return on;
}

std::atomic<bool> g_aborted{false};
std::mutex g_status_mu;
std::string g_status;
thread_local std::string t_status;
thread_local bool t_rows_bad = false;

std::string error_text() {
    const char *e = mb3d_hip_last_error();
    return e ? std::string(e) : std::string();
}

struct CallCtx {
    const Progress *progress;
    const Cancel *cancel;
};

// Marches rows [row0, row0 + rows) into `words` (5 per pixel, host memory).
int march_rows(const mb3d_hip_params &p, const mb3d_hip_post_job *post, size_t row0, size_t rows, uint32_t *words,
               const mb3d_hip_options *options) {
    mb3d_hip_out out;
    std::memset(&out, 0, sizeof out);
    out.gbuf = words;
    return mb3d_hip_march(&p, post, (int32_t)row0, (int32_t)rows, &out, options, nullptr, nullptr, nullptr);
}

// The start of a card calculation: switched on, not stopped, supported,
// device open.  An empty text means go.
std::string may_start(const mb3d_hip_params &p) {
    if (!enabled()) return "the graphics card is switched off";
    if (aborted()) return "stopped";
    if (auto why = unsupported(p)) return *why;
    if (mb3d_hip_init(0) != MB3D_HIP_OK) return error_text();
    return std::string();
}

}  // namespace

SiLight unpack(const uint32_t w[5]) {
    SiLight s;
    s.normal[0] = (int16_t)(uint16_t)(w[0] & 0xFFFF);
    s.normal[1] = (int16_t)(uint16_t)(w[0] >> 16);
    s.normal[2] = (int16_t)(uint16_t)(w[1] & 0xFFFF);
    s.shadow = (uint16_t)(w[1] >> 16);
    s.zpos_fine = w[2];
    s.si_gradient = (uint16_t)(w[3] & 0xFFFF);
    s.otrap = (uint16_t)(w[3] >> 16);
    s.amb_shadow = (uint16_t)(w[4] & 0xFFFF);
    return s;
}

void pack(const SiLight &s, uint32_t w[5]) {
    w[0] = (uint32_t)(uint16_t)s.normal[0] | ((uint32_t)(uint16_t)s.normal[1] << 16);
    w[1] = (uint32_t)(uint16_t)s.normal[2] | ((uint32_t)s.shadow << 16);
    w[2] = s.zpos_fine;
    w[3] = (uint32_t)s.si_gradient | ((uint32_t)s.otrap << 16);
    w[4] = s.amb_shadow;
}

bool enabled() { return enabled_flag().load(); }
void set_enabled(bool on) { enabled_flag().store(on); }

std::string last_status() {
    std::lock_guard<std::mutex> lk(g_status_mu);
    return g_status;
}

std::string thread_status() { return t_status; }

void set_status(const std::string &s) {
    t_status = s;
    std::lock_guard<std::mutex> lk(g_status_mu);
    g_status = s;
}

bool rows_overflowed() { return t_rows_bad; }
void clear_rows_overflowed() { t_rows_bad = false; }
void abort_all() { g_aborted.store(true); }
bool aborted() { return g_aborted.load(); }

std::optional<std::string> unsupported(const mb3d_hip_params &p) {
    char why[512] = {0};
    const int st = mb3d_hip_supported(&p, why, sizeof why);
    if (st == MB3D_HIP_OK) return std::nullopt;
    if (st < 0) return error_text();
    return std::string(why);
}

std::optional<std::vector<SiLight>> march(const mb3d_hip_params &p, const mb3d_hip_post_job *post,
                                          const Progress &progress, const Cancel &cancel, const Sink &sink,
                                          const mb3d_hip_options *options) {
    t_rows_bad = false;
    const std::string why = may_start(p);
    if (!why.empty()) {
        set_status("CPU: " + why);
        return std::nullopt;
    }
    const size_t w = (size_t)p.width, h = (size_t)p.height;
    std::vector<uint32_t> words(w * h * 5);
    std::vector<SiLight> gbuf(w * h);
    for (size_t row0 = 0; row0 < h; row0 += BAND_ROWS) {
        if (cancel && cancel()) {
            set_status("CPU: cancelled");
            return std::nullopt;
        }
        const size_t rows = std::min(BAND_ROWS, h - row0);
        const int st = march_rows(p, post, row0, rows, words.data() + row0 * w * 5, options);
        if (st == MB3D_HIP_OVERFLOW) {
            t_rows_bad = true;
            set_status("CPU: single precision overflowed");
            return std::nullopt;
        }
        if (st != MB3D_HIP_OK) {
            set_status("CPU: " + error_text());
            return std::nullopt;
        }
        for (size_t i = row0 * w; i < (row0 + rows) * w; i++) gbuf[i] = unpack(&words[i * 5]);
        if (sink)
            for (size_t y = row0; y < row0 + rows; y++) sink(y, gbuf.data() + y * w, w);
        if (progress) progress(row0 + rows, h);
    }
    set_status("GPU (MI355X, HIP)");
    return gbuf;
}

Status march_bands(const mb3d_hip_params &p, const mb3d_hip_post_job *post, const Bands &bands,
                   const Cancel &cancel, const Sink &sink, const mb3d_hip_options *options) {
    t_rows_bad = false;
    const std::string why = may_start(p);
    if (!why.empty()) return Status::fail(why);
    const size_t w = (size_t)p.width, h = (size_t)p.height;
    std::vector<uint32_t> words;
    std::vector<SiLight> row(w);
    for (size_t i = 0;; i++) {
        if (cancel && cancel()) return Status::fail("cancelled");
        const auto b = bands(i);
        if (!b) break;
        const size_t row0 = b->first, rows = b->second;
        if (row0 >= h || rows == 0) continue;
        const size_t n = std::min(rows, h - row0);
        words.resize(n * w * 5);
        const int st = march_rows(p, post, row0, n, words.data(), options);
        if (st == MB3D_HIP_OVERFLOW) {
            t_rows_bad = true;
            return Status::fail("single precision overflowed");
        }
        if (st != MB3D_HIP_OK) return Status::fail(error_text());
        for (size_t y = 0; y < n; y++) {
            for (size_t x = 0; x < w; x++) row[x] = unpack(&words[(y * w + x) * 5]);
            if (sink) sink(row0 + y, row.data(), w);
        }
    }
    return Status{};
}

Status post_on_card(const mb3d_hip_params &p, std::vector<SiLight> &gbuf, const mb3d_hip_post_job &post,
                    const Progress &progress, const Cancel &cancel) {
    if (!enabled()) return Status::fail("the graphics card is switched off");
    if (auto why = unsupported(p)) return Status::fail(*why);
    if (mb3d_hip_init(0) != MB3D_HIP_OK) return Status::fail(error_text());
    const size_t w = (size_t)p.width, h = (size_t)p.height;
    if (gbuf.size() < w * h) return Status::fail("the G-buffer is smaller than the picture");
    std::vector<uint32_t> words;
    for (size_t row0 = 0; row0 < h; row0 += BAND_ROWS) {
        if (cancel && cancel()) return Status::fail("cancelled");
        const size_t rows = std::min(BAND_ROWS, h - row0);
        words.resize(rows * w * 5);
        for (size_t i = 0; i < rows * w; i++) pack(gbuf[row0 * w + i], &words[i * 5]);
        const int st = mb3d_hip_post(&p, &post, (int32_t)row0, (int32_t)rows, words.data(), 0, nullptr);
        if (st != MB3D_HIP_OK) return Status::fail(error_text());
        for (size_t i = 0; i < rows * w; i++) gbuf[row0 * w + i] = unpack(&words[i * 5]);
        if (progress) progress(row0 + rows, h);
    }
    return Status{};
}

namespace {

// Which rows are where (render::Board, src/render.rs:1380-1470, one card):
// delivered rows, rows a second worker backs up, and the card's bands.
class Board {
  public:
    explicit Board(size_t h) : delivered_(h), backup_(h) {
        for (size_t y = 0; y < h; y++) delivered_[y] = backup_[y] = false;
    }
    void took(size_t r0, size_t n) {
        std::lock_guard<std::mutex> lk(mu_);
        card_.push_back({r0, n});
    }
    // the first result of row y counts
    bool deliver(size_t y) { return !delivered_[y].exchange(true); }
    bool is_delivered(size_t y) const { return delivered_[y].load(); }
    // up to `max` rows in a row the other side took and has not delivered,
    // nobody else backing them up (the card steals from the CPU's rows, a
    // CPU thread from the card's bands)
    std::optional<std::pair<size_t, size_t>> steal(bool for_card, size_t max, const std::vector<size_t> &cpu_taken) {
        std::vector<std::pair<size_t, size_t>> from;
        if (for_card) {
            for (size_t y : cpu_taken) from.push_back({y, 1});
        } else {
            std::lock_guard<std::mutex> lk(mu_);
            from = card_;
        }
        auto open = [&](size_t y) { return !delivered_[y].load() && !backup_[y].load(); };
        for (auto [r0, n] : from) {
            size_t y = r0;
            while (y < r0 + n && !open(y)) y++;
            if (y >= r0 + n) continue;
            size_t m = 0;
            while (y + m < r0 + n && m < std::max<size_t>(max, 1) && open(y + m) && !backup_[y + m].exchange(true)) m++;
            if (m > 0) {
                if (for_card) took(y, m);
                return std::make_pair(y, m);
            }
        }
        return std::nullopt;
    }
    // every row of the card's bands is delivered (by it or by others)
    bool card_done() {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto [r0, n] : card_)
            for (size_t y = r0; y < r0 + n; y++)
                if (!delivered_[y].load()) return false;
        return true;
    }
    std::vector<std::pair<size_t, size_t>> card_bands() {
        std::lock_guard<std::mutex> lk(mu_);
        return card_;
    }

  private:
    std::vector<std::atomic<bool>> delivered_, backup_;
    std::mutex mu_;
    std::vector<std::pair<size_t, size_t>> card_;
};

}  // namespace

Mixed march_mixed(const mb3d_hip_params &p, const mb3d_hip_post_job *post, size_t threads, const CpuRow &cpu_row,
                  const Sink &sink, const Cancel &cancel, const Progress &progress, const Done &done,
                  const std::vector<SiLight> *have, size_t band_rows, const mb3d_hip_options *options) {
    const size_t w = (size_t)std::max(p.width, 1), h = (size_t)std::max(p.height, 0);
    Mixed res;
    res.gbuf.assign(w * h, SiLight{});
    auto is_done = [&](size_t y) { return done && done(y); };
    if (have)
        for (size_t y = 0; y < h; y++)
            if (is_done(y) && have->size() >= (y + 1) * w) std::copy_n(have->data() + y * w, w, res.gbuf.data() + y * w);
    Board board(h);
    std::mutex free_mu, put_mu;
    size_t top = 0, bottom = h;  // the free rows [top, bottom)
    std::vector<size_t> cpu_taken;
    std::atomic<size_t> done_rows{0}, card_rows{0}, cpu_rows{0};
    for (size_t y = 0; y < h; y++)
        if (is_done(y)) {
            board.deliver(y);
            done_rows++;
        }
    auto stopped = [&] { return cancel && cancel(); };
    auto left = [&] {
        std::lock_guard<std::mutex> lk(free_mu);
        while (top < bottom && is_done(top)) top++;
        while (bottom > top && is_done(bottom - 1)) bottom--;
        return top < bottom;
    };
    auto put = [&](size_t y, const SiLight *row) {
        std::copy_n(row, w, res.gbuf.data() + y * w);
        std::lock_guard<std::mutex> lk(put_mu);
        if (sink) sink(y, row, w);
        const size_t d = ++done_rows;
        if (progress) progress(d, h);
    };
    // the card: bands from the top, then rows backed up from the CPU's
    std::mutex pending_mu;
    std::vector<size_t> pending;  // the card's delivered rows, until its verdict
    Status card_status;
    bool card_joined = false;
    const std::string why = may_start(p);
    std::thread card;
    if (why.empty()) {
        card_joined = true;
        card = std::thread([&] {
            Bands bands = [&](size_t) -> std::optional<std::pair<size_t, size_t>> {
                {
                    std::lock_guard<std::mutex> lk(free_mu);
                    while (top < bottom && is_done(top)) top++;
                    if (top < bottom) {
                        size_t n = 0;
                        while (top + n < bottom && n < band_rows && !is_done(top + n)) n++;
                        const size_t r0 = top;
                        top += n;
                        board.took(r0, n);
                        return std::make_pair(r0, n);
                    }
                }
                std::vector<size_t> taken;
                {
                    std::lock_guard<std::mutex> lk(free_mu);
                    taken = cpu_taken;
                }
                return board.steal(true, band_rows, taken);
            };
            Cancel stop = [&] { return stopped() || (!left() && board.card_done()); };
            Sink card_sink = [&](size_t y, const SiLight *row, size_t) {
                if (!board.deliver(y)) return;
                {
                    std::lock_guard<std::mutex> lk(pending_mu);
                    pending.push_back(y);
                }
                card_rows++;
                put(y, row);
            };
            card_status = march_bands(p, post, bands, stop, card_sink, options);
        });
    }
    // the CPU's threads: rows from the bottom, then rows backed up from the card's bands
    auto worker = [&] {
        std::vector<SiLight> row(w);
        for (;;) {
            if (stopped()) return;
            std::optional<size_t> y;
            {
                std::lock_guard<std::mutex> lk(free_mu);
                while (bottom > top && is_done(bottom - 1)) bottom--;
                if (bottom > top) {
                    y = --bottom;
                    cpu_taken.push_back(*y);
                }
            }
            if (!y) {
                auto r = board.steal(false, 1, {});
                if (!r) {
                    // nothing free and nothing to back up: wait for the card's open rows
                    if (!card_joined || board.card_done()) return;
                    std::this_thread::sleep_for(std::chrono::milliseconds(1));
                    continue;
                }
                y = r->first;
            }
            cpu_row(*y, row.data());
            if (board.deliver(*y)) {
                cpu_rows++;
                put(*y, row.data());
            }
        }
    };
    std::vector<std::thread> pool;
    for (size_t i = 0; i < std::max<size_t>(threads, 1); i++) pool.emplace_back(worker);
    for (auto &t : pool) t.join();
    if (card.joinable()) card.join();
    // a card that overflowed: its rows are redone here (the first result did
    // not count after all)
    if (card_joined && !card_status.ok && card_status.error != "cancelled" && !stopped()) {
        std::vector<size_t> redo;
        {
            std::lock_guard<std::mutex> lk(pending_mu);
            redo = pending;
        }
        std::vector<SiLight> row(w);
        for (size_t y : redo) {
            cpu_row(y, row.data());
            put(y, row.data());
            cpu_rows++;
        }
        card_rows = 0;
    }
    // rows a stopped card took and never delivered, with the CPU's threads gone
    for (size_t y = 0; y < h && !stopped(); y++) {
        if (board.is_delivered(y)) continue;
        std::vector<SiLight> row(w);
        cpu_row(y, row.data());
        if (board.deliver(y)) {
            cpu_rows++;
            put(y, row.data());
        }
    }
    res.card_rows = card_rows.load();
    res.cpu_rows = cpu_rows.load();
    if (stopped()) res.status = Status::fail("cancelled");
    set_status(card_joined && res.card_rows ? "GPU (MI355X, HIP): the march shared with the CPU"
                                             : "CPU: " + (why.empty() ? card_status.error : why));
    return res;
}

McShared mc_batch_shared(const mb3d_hip_params &p, const std::vector<mb3d_hip_mc_task> &tasks, bool secant,
                         size_t threads, const CpuTasks &cpu, size_t chunk, size_t cpu_chunk, const Cancel &cancel) {
    const size_t n = tasks.size();
    McShared out;
    out.res.assign(n, mb3d_hip_mc_res{});
    out.side.assign(n, 0);
    std::mutex mu;
    size_t front = 0, back = n;  // the free tasks [front, back)
    std::atomic<size_t> card_n{0}, cpu_n{0};
    std::atomic<bool> stop{false};
    auto stopped = [&] {
        if (!stop.load() && cancel && cancel()) stop = true;
        return stop.load();
    };
    // up to `max` free tasks [a, b) from the front (the card) or the back (the CPU)
    auto claim = [&](bool from_front, size_t max, size_t &a, size_t &b) {
        std::lock_guard<std::mutex> lk(mu);
        if (front >= back) return false;
        const size_t m = std::min(std::max<size_t>(max, 1), back - front);
        if (from_front) {
            a = front;
            b = front = front + m;
        } else {
            b = back;
            a = back = back - m;
        }
        return true;
    };
    auto on_cpu = [&](size_t a, size_t b) {
        cpu(tasks.data() + a, b - a, out.res.data() + a);
        std::fill(out.side.begin() + a, out.side.begin() + b, (uint8_t)2);
        cpu_n += b - a;
    };
    std::vector<std::thread> pool;
    if (cpu)
        for (size_t i = 0; i < threads; i++)
            pool.emplace_back([&] {
                size_t a = 0, b = 0;
                while (!stopped() && claim(false, cpu_chunk, a, b)) on_cpu(a, b);
            });
    Status card_status;
    const std::string why = may_start(p);
    if (!why.empty()) {
        card_status = Status::fail(why);
    } else {
        size_t a = 0, b = 0;
        while (!stopped() && claim(true, chunk, a, b)) {
            if (mb3d_hip_mc_batch(&p, tasks.data() + a, (int64_t)(b - a), secant ? 1 : 0, out.res.data() + a, 0) !=
                MB3D_HIP_OK) {
                card_status = Status::fail(error_text());
                if (cpu) on_cpu(a, b);  // the failed chunk goes to the CPU marcher
                break;
            }
            std::fill(out.side.begin() + a, out.side.begin() + b, (uint8_t)1);
            card_n += b - a;
        }
    }
    // a card that did not start or failed: the calling thread joins the CPU's side
    if (!card_status.ok && cpu) {
        size_t a = 0, b = 0;
        while (!stopped() && claim(false, cpu_chunk, a, b)) on_cpu(a, b);
    }
    for (auto &t : pool) t.join();
    out.card = card_n;
    out.cpu = cpu_n;
    if (stop.load() && out.card + out.cpu < n)
        out.status = Status::fail("cancelled");
    else if (out.card + out.cpu < n)
        out.status = Status::fail(card_status.ok ? std::string("no CPU marcher") : card_status.error);
    else
        out.status = card_status;
    return out;
}

}  // namespace mb3d::gpu
