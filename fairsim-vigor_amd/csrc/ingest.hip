// Live stream ingest: the camera wire format of transport/ImageWrapper.java
// and the SIM sequence assembly of livemode/SimSequenceExtractor.java,
// restated as a push-driven state machine (host code).
//
// The reference runs one blocking thread per channel whose loop pulls frames
// through getSorted(); here each pushed frame advances the same loop until
// it would block (no frame available), and the loop state (sorter, sync
// search, sequence being collected) is kept between pushes.  With a server
// attached, assembled sequences go to fsim_server_submit through pinned
// staging, and fsim_ingest_result hands back reconstructions in submission
// order.
#include "fsim_internal.h"

#include <climits>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "the wire format is little endian, as this host");

namespace fsim {
namespace {

template <class T>
T rd(const uint8_t *p) {
    T v;
    std::memcpy(&v, p, sizeof(T));
    return v;
}
template <class T>
void wr(uint8_t *p, T v) {
    std::memcpy(p, &v, sizeof(T));
}

constexpr int kHeader = 128;
constexpr int kSyncJitter = 50;   // PerChannelBuffer.SYNCJITTER

// ImageWrapper.parseHeader (ImageWrapper.java:443-476)
int parse_header(const uint8_t *h, int maxw, int maxh, fsim_frame_header &o) {
    o.seq_nr = rd<int64_t>(h + 0);
    o.npix = rd<int32_t>(h + 12);
    o.version = (int8_t)h[16];
    o.bpp = (int8_t)h[17];
    o.width = rd<int16_t>(h + 32);
    o.height = rd<int16_t>(h + 34);
    o.pos_a = rd<int16_t>(h + 36);
    o.pos_b = rd<int16_t>(h + 38);
    o.pos0 = rd<int16_t>(h + 40);
    o.pos1 = rd<int16_t>(h + 42);
    o.pos2 = rd<int32_t>(h + 44);
    o.time_camera = rd<int64_t>(h + 48);
    o.time_capture = rd<int64_t>(h + 56);
    o.time_record = rd<int64_t>(h + 64);
    if (o.width > maxw || o.height > maxh)
        return fail(FSIM_ERR_ARG, "wire header: image too large, w " + std::to_string(o.width) + "/" +
                                      std::to_string(maxw) + ", h " + std::to_string(o.height) + "/" +
                                      std::to_string(maxh));
    if (o.npix != o.width * o.height) return fail(FSIM_ERR_ARG, "wire header: Width x Height != Len");
    if (o.bpp != 1 && o.bpp != 2) return fail(FSIM_ERR_ARG, "wire header: bpp neither 1 nor 2");
    if (o.version != 1) return fail(FSIM_ERR_ARG, "wire header: unsupported version");
    if (o.width <= 0 || o.height <= 0) return fail(FSIM_ERR_ARG, "wire header: empty image");
    return FSIM_OK;
}

// ImageWrapper.writeHeader (ImageWrapper.java:420-441)
void write_header(const fsim_frame_header &o, uint8_t *h) {
    std::memset(h, 0, kHeader);
    wr<int64_t>(h + 0, o.seq_nr);
    wr<int32_t>(h + 12, o.width * o.height);
    h[16] = 1;
    h[17] = (uint8_t)o.bpp;
    wr<int16_t>(h + 32, (int16_t)o.width);
    wr<int16_t>(h + 34, (int16_t)o.height);
    wr<int16_t>(h + 36, (int16_t)o.pos_a);
    wr<int16_t>(h + 38, (int16_t)o.pos_b);
    wr<int16_t>(h + 40, (int16_t)o.pos0);
    wr<int16_t>(h + 42, (int16_t)o.pos1);
    wr<int32_t>(h + 44, o.pos2);
    wr<int64_t>(h + 48, o.time_camera);
    wr<int64_t>(h + 56, o.time_capture);
    wr<int64_t>(h + 64, o.time_record);
}

struct Frame {
    fsim_frame_header h;
    std::vector<uint16_t> px;   // ImageWrapper.getPixels as 16-bit patterns
};

struct Seq {
    fsim_ingest_seq_info info;
    std::vector<uint16_t> px;   // ndir*nphase frames
    bool ok = true;             // every frame had the first frame's size
};

enum class St { Search, Skip, Collect };

struct Chan {
    int number = 0;
    // sorter (PerChannelBuffer.getSorted / sortFromQueue / setSeqNr)
    std::deque<Frame *> raw;
    std::map<int64_t, Frame *> sortbuf;
    int64_t seqnr = LLONG_MAX;
    bool resync = true;   // setSeqNr pending
    // sequence detection (PerChannelBuffer.run)
    St st = St::Search;
    int counter = 0, k = 0, i = 0;
    int64_t last_ts = 0;
    Seq *cur = nullptr;
    std::deque<Seq *> done;
    fsim_ingest_stats stats{};
};

}  // namespace
}  // namespace fsim

using namespace fsim;

struct fsim_ingest_s {
    fsim_ingest_config cfg;
    int raws = 0;
    std::vector<Chan> ch;
    std::map<int, int> index;   // channel number -> index
    std::vector<Frame *> frame_pool;
    std::vector<Seq *> seq_pool;
    std::mutex mu;
    // server feed
    fsim_server srv = nullptr;
    std::vector<int> srv_ch;
    int n = 0, rr = 0;
    std::vector<uint16_t *> raw_pin;
    std::vector<float *> out_pin;
    std::vector<int> free_slots;
    struct Inflight {
        uint64_t ticket;
        int slot;
        fsim_ingest_seq_info info;
    };
    std::deque<Inflight> inflight;
    bool waiting = false;

    Frame *new_frame() {
        if (frame_pool.empty()) return new Frame();
        Frame *f = frame_pool.back();
        frame_pool.pop_back();
        return f;
    }
    void free_frame(Frame *f) { frame_pool.push_back(f); }
    Seq *new_seq() {
        Seq *s;
        if (seq_pool.empty()) {
            s = new Seq();
        } else {
            s = seq_pool.back();
            seq_pool.pop_back();
        }
        s->ok = true;
        return s;
    }
    void free_seq(Seq *s) { seq_pool.push_back(s); }

    // sort buffer insert; a second frame with the same number is an error
    // (SortBuffer.add, SimSequenceExtractor.java:488-499)
    int sort_add(Chan &c, Frame *f) {
        auto r = c.sortbuf.emplace(f->h.seq_nr, f);
        if (!r.second) {
            free_frame(f);
            return fail(FSIM_ERR_ARG, "ingest: same seqNr " + std::to_string(f->h.seq_nr) + " twice");
        }
        return FSIM_OK;
    }

    void sort_clear(Chan &c) {
        for (auto &e : c.sortbuf) free_frame(e.second);
        c.sortbuf.clear();
    }

    // getSorted: the next frame in sequence-number order, or nullptr when the
    // reference would block for input (SimSequenceExtractor.java:262-330)
    int next(Chan &c, Frame *&out) {
        out = nullptr;
        for (;;) {
            if (c.resync) {
                // setSeqNr: the lowest number among the next ndir*nphase frames
                while ((int)c.sortbuf.size() < raws && !c.raw.empty()) {
                    Frame *f = c.raw.front();
                    c.raw.pop_front();
                    const int64_t sn = f->h.seq_nr;
                    int rc = sort_add(c, f);
                    if (rc) return rc;
                    if (sn < c.seqnr) c.seqnr = sn;
                }
                if ((int)c.sortbuf.size() < raws) return FSIM_OK;
                c.resync = false;
            }
            if (!c.sortbuf.empty()) {
                auto it = c.sortbuf.find(c.seqnr);
                if (it != c.sortbuf.end()) {
                    out = it->second;
                    c.sortbuf.erase(it);
                    c.seqnr++;
                    return FSIM_OK;
                }
            }
            if (c.raw.empty()) return FSIM_OK;
            Frame *f = c.raw.front();
            c.raw.pop_front();
            if (f->h.seq_nr == c.seqnr) {
                c.seqnr++;
                out = f;
                return FSIM_OK;
            }
            if ((int)c.sortbuf.size() >= cfg.sort_capacity) {
                // overflow: the frame is lost, resynchronise
                free_frame(f);
                c.stats.dropped_frames++;
                c.seqnr++;
                sort_clear(c);
                c.seqnr = LLONG_MAX;
                c.resync = true;
                continue;
            }
            int rc = sort_add(c, f);
            if (rc) return rc;
        }
    }

    // MTool.avr_ushort (linalg/MTool.java:68-76)
    static double avr_ushort(const Frame *f) {
        double s = 0;
        for (uint16_t v : f->px) s += v;
        return s / f->px.size();
    }

    void seq_frame(Chan &c, Frame *f) {
        if (c.i == 0) {
            c.cur = new_seq();
            c.cur->info.channel = (int)(&c - ch.data());
            c.cur->info.width = f->h.width;
            c.cur->info.height = f->h.height;
            c.cur->info.first_seq_nr = f->h.seq_nr;
            c.cur->info.time_camera = f->h.time_camera;
            c.cur->info.pos2 = f->h.pos2;
            c.cur->px.resize((size_t)raws * f->px.size());
        }
        Seq *s = c.cur;
        const size_t fp = (size_t)s->info.width * s->info.height;
        if (f->h.width != s->info.width || f->h.height != s->info.height)
            s->ok = false;
        else
            std::memcpy(s->px.data() + (size_t)c.i * fp, f->px.data(), fp * sizeof(uint16_t));
        c.last_ts = f->h.time_camera;
        if (++c.i == raws) {
            c.i = 0;
            c.cur = nullptr;
            if (!s->ok || (int)c.done.size() >= cfg.queue_capacity) {
                c.stats.missed_sequences++;
                free_seq(s);
            } else {
                c.done.push_back(s);
                c.stats.sequences++;
            }
            if (++c.k == cfg.seq_count) {
                c.k = 0;
                c.st = St::Search;
                c.counter = 0;
            }
        }
    }

    // PerChannelBuffer.run (SimSequenceExtractor.java:369-470), advanced
    // until it needs another frame
    int advance(Chan &c) {
        for (;;) {
            Frame *f;
            int rc = next(c, f);
            if (rc) return rc;
            if (!f) return FSIM_OK;
            switch (c.st) {
                case St::Search: {
                    const int64_t cur = f->h.time_camera;
                    if (cfg.sync_mode == 0) {
                        if (cur - c.last_ts > cfg.sync_delay) {
                            // the frame after the gap starts the sequence: back into the sorter
                            if ((int)c.sortbuf.size() >= cfg.sort_capacity) {
                                free_frame(f);
                                return fail(FSIM_ERR_STATE, "ingest: sort buffer overflow at a sync frame");
                            }
                            rc = sort_add(c, f);
                            if (rc) return rc;
                            c.seqnr--;
                            c.stats.sync_frames++;
                            c.st = St::Collect;
                            c.k = c.i = 0;
                            continue;
                        }
                    } else {
                        if (std::llabs(cur - c.last_ts - cfg.sync_delay) < kSyncJitter) {
                            c.stats.sync_frames++;
                            c.st = St::Collect;
                            c.k = c.i = 0;
                            free_frame(f);
                            continue;
                        }
                        if (avr_ushort(f) > cfg.sync_avr) {
                            c.stats.sync_frames++;
                            c.st = St::Skip;
                            free_frame(f);
                            continue;
                        }
                    }
                    c.last_ts = cur;
                    if (++c.counter >= raws * cfg.seq_count) {
                        c.stats.no_sync_windows++;
                        c.counter = 0;
                    }
                    free_frame(f);
                    break;
                }
                case St::Skip:
                    free_frame(f);
                    c.st = St::Collect;
                    c.k = c.i = 0;
                    break;
                case St::Collect:
                    seq_frame(c, f);
                    free_frame(f);
                    break;
            }
        }
    }

    // submit waiting sequences while staging slots are free; channels in
    // turn, as the reference joins one sequence per channel
    int pump() {
        if (!srv) return FSIM_OK;
        const int nc = (int)ch.size();
        while (!free_slots.empty()) {
            int pick = -1;
            for (int t = 0; t < nc; t++) {
                const int c = (rr + t) % nc;
                if (!ch[c].done.empty()) {
                    pick = c;
                    break;
                }
            }
            if (pick < 0) return FSIM_OK;
            Chan &c = ch[pick];
            Seq *s = c.done.front();
            if (s->info.width != n || s->info.height != n) {
                c.done.pop_front();
                c.stats.missed_sequences++;
                free_seq(s);
                continue;
            }
            const int slot = free_slots.back();
            std::memcpy(raw_pin[slot], s->px.data(), s->px.size() * sizeof(uint16_t));
            uint64_t ticket = 0;
            int rc = fsim_server_submit(srv, srv_ch[pick], raw_pin[slot], out_pin[slot], 0, &ticket);
            if (rc == FSIM_ERR_BUSY) return FSIM_OK;   // server queue full: keep it, retry later
            c.done.pop_front();
            rr = (pick + 1) % nc;
            if (rc) {
                c.stats.missed_sequences++;
                free_seq(s);
                return rc;
            }
            free_slots.pop_back();
            inflight.push_back({ticket, slot, s->info});
            free_seq(s);
        }
        return FSIM_OK;
    }

    void release_server() {
        for (auto p : raw_pin) (void)hipHostFree(p);
        for (auto p : out_pin) (void)hipHostFree(p);
        raw_pin.clear();
        out_pin.clear();
        free_slots.clear();
        srv = nullptr;
    }

    ~fsim_ingest_s() {
        for (auto &e : inflight) (void)fsim_server_wait(srv, e.ticket, -1, nullptr);
        release_server();
        for (auto &c : ch) {
            for (Frame *f : c.raw) delete f;
            for (auto &e : c.sortbuf) delete e.second;
            for (Seq *s : c.done) delete s;
            delete c.cur;
        }
        for (Frame *f : frame_pool) delete f;
        for (Seq *s : seq_pool) delete s;
    }
};

extern "C" {

int fsim_wire_parse_header(const uint8_t *hdr, int max_width, int max_height, fsim_frame_header *out) {
    if (!hdr || !out) return fail(FSIM_ERR_ARG, "wire header: null argument");
    return parse_header(hdr, max_width, max_height, *out);
}

int fsim_wire_write_header(const fsim_frame_header *h, uint8_t *hdr) {
    if (!h || !hdr) return fail(FSIM_ERR_ARG, "wire header: null argument");
    if (h->bpp != 1 && h->bpp != 2) return fail(FSIM_ERR_ARG, "wire header: bpp neither 1 nor 2");
    write_header(*h, hdr);
    return FSIM_OK;
}

int fsim_ingest_create(const fsim_ingest_config *cfg, const int *numbers, int nchannels, fsim_ingest *out) {
    if (!cfg || !numbers || !out || nchannels < 1) return fail(FSIM_ERR_ARG, "ingest_create: bad arguments");
    if (cfg->ndir < 1 || cfg->nphase < 1 || cfg->seq_count < 1 || cfg->sort_capacity < cfg->ndir * cfg->nphase + 1 ||
        cfg->queue_capacity < 1 || cfg->max_width < 1 || cfg->max_height < 1)
        return fail(FSIM_ERR_ARG, "ingest_create: bad configuration (sort buffer must exceed ndir*nphase)");
    if (cfg->sync_mode != 0 && cfg->sync_mode != 1)
        return fail(FSIM_ERR_ARG, "ingest_create: missmatch in syncMode, expected 0 or 1");
    fsim_ingest_s *in = new fsim_ingest_s();
    in->cfg = *cfg;
    in->raws = cfg->ndir * cfg->nphase;
    in->ch.resize(nchannels);
    for (int c = 0; c < nchannels; c++) {
        in->ch[c].number = numbers[c];
        if (!in->index.emplace(numbers[c], c).second) {
            delete in;
            return fail(FSIM_ERR_ARG, "ingest_create: channel number used twice");
        }
    }
    *out = in;
    return FSIM_OK;
}

int fsim_ingest_destroy(fsim_ingest in) {
    delete in;
    return FSIM_OK;
}

int fsim_ingest_push(fsim_ingest in, const uint8_t *packet, int64_t len) {
    if (!in || !packet || len < kHeader) return fail(FSIM_ERR_ARG, "ingest_push: short packet");
    std::lock_guard<std::mutex> lk(in->mu);
    fsim_frame_header h;
    int rc = parse_header(packet, in->cfg.max_width, in->cfg.max_height, h);
    if (rc) return rc;
    const int64_t np = (int64_t)h.width * h.height;
    if (len < kHeader + np * h.bpp) return fail(FSIM_ERR_ARG, "ingest_push: payload shorter than the header says");
    auto it = in->index.find(h.pos1);
    if (it == in->index.end())
        return fail(FSIM_ERR_ARG, "ingest_push: received data packet w/o channel: " + std::to_string(h.pos1));
    Chan &c = in->ch[it->second];
    Frame *f = in->new_frame();
    f->h = h;
    f->px.resize(np);
    const uint8_t *p = packet + kHeader;
    if (h.bpp == 2) {
        std::memcpy(f->px.data(), p, np * 2);
    } else {
        // ImageWrapper.getPixels: ret[i] = buffer[i + HEADERSIZE], a signed byte widened to short
        for (int64_t i = 0; i < np; i++) f->px[i] = (uint16_t)(int16_t)(int8_t)p[i];
    }
    c.stats.frames++;
    c.raw.push_back(f);
    rc = in->advance(c);
    if (rc) return rc;
    return in->pump();
}

int fsim_ingest_pop(fsim_ingest in, int chi, uint16_t *out, int64_t max_pixels, fsim_ingest_seq_info *info) {
    if (!in || !out) return fail(FSIM_ERR_ARG, "ingest_pop: null argument");
    std::lock_guard<std::mutex> lk(in->mu);
    if (chi < 0 || chi >= (int)in->ch.size()) return fail(FSIM_ERR_ARG, "ingest_pop: channel index out of range");
    Chan &c = in->ch[chi];
    if (c.done.empty()) return 0;
    Seq *s = c.done.front();
    if ((int64_t)s->px.size() > max_pixels) return fail(FSIM_ERR_ARG, "ingest_pop: output buffer too small");
    std::memcpy(out, s->px.data(), s->px.size() * sizeof(uint16_t));
    if (info) *info = s->info;
    c.done.pop_front();
    in->free_seq(s);
    return 1;
}

int fsim_ingest_get_stats(fsim_ingest in, int chi, fsim_ingest_stats *st) {
    if (!in || !st) return fail(FSIM_ERR_ARG, "ingest_stats: null argument");
    std::lock_guard<std::mutex> lk(in->mu);
    if (chi < 0 || chi >= (int)in->ch.size()) return fail(FSIM_ERR_ARG, "ingest_stats: channel index out of range");
    *st = in->ch[chi].stats;
    return FSIM_OK;
}

int fsim_ingest_attach_server(fsim_ingest in, fsim_server srv, const int *server_channels, int n, int max_inflight) {
    if (!in || !srv || !server_channels || n < 1 || max_inflight < 1)
        return fail(FSIM_ERR_ARG, "ingest_attach_server: bad arguments");
    int rc = ensure_init();
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(in->mu);
    if (in->srv) return fail(FSIM_ERR_STATE, "ingest_attach_server: a server is attached already");
    in->srv_ch.assign(server_channels, server_channels + in->ch.size());
    in->n = n;
    const size_t rawb = (size_t)in->raws * n * n * sizeof(uint16_t), outb = (size_t)4 * n * n * sizeof(float);
    for (int i = 0; i < max_inflight; i++) {
        void *a = nullptr, *b = nullptr;
        if (hipHostMalloc(&a, rawb) != hipSuccess || hipHostMalloc(&b, outb) != hipSuccess) {
            (void)hipHostFree(a);
            in->release_server();
            return fail(FSIM_ERR_OOM, "ingest_attach_server: pinned staging allocation failed");
        }
        in->raw_pin.push_back((uint16_t *)a);
        in->out_pin.push_back((float *)b);
        in->free_slots.push_back(i);
    }
    in->srv = srv;
    return in->pump();
}

int fsim_ingest_result(fsim_ingest in, int timeout_ms, float *out, fsim_ingest_seq_info *info) {
    if (!in || !out) return fail(FSIM_ERR_ARG, "ingest_result: null argument");
    fsim_ingest_s::Inflight e;
    {
        std::lock_guard<std::mutex> lk(in->mu);
        if (!in->srv) return fail(FSIM_ERR_STATE, "ingest_result: no server attached");
        if (in->waiting) return fail(FSIM_ERR_STATE, "ingest_result: one caller at a time");
        int rc = in->pump();
        if (rc) return rc;
        if (in->inflight.empty()) return fail(FSIM_ERR_BUSY, "ingest_result: no sequence in flight");
        e = in->inflight.front();
        in->waiting = true;
    }
    int rc = fsim_server_wait(in->srv, e.ticket, timeout_ms, nullptr);
    std::lock_guard<std::mutex> lk(in->mu);
    in->waiting = false;
    if (rc == FSIM_ERR_BUSY) return rc;   // timed out: still in flight
    in->inflight.pop_front();
    if (rc == FSIM_OK) {
        std::memcpy(out, in->out_pin[e.slot], (size_t)4 * in->n * in->n * sizeof(float));
        if (info) *info = e.info;
    }
    in->free_slots.push_back(e.slot);
    const int rp = in->pump();
    return rc ? rc : rp;
}

}  // extern "C"
