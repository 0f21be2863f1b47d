// ffvship_amd.cpp - throughput driver for libvship_amd.so over raw planar
// YUV files: the FFVship pipeline (reference src/FFVship.cpp:43-145,
// 385-444, src/ffvship_utility/ffmpegmain.hpp:40-172) with the same frame
// selection options, JSON and live-score formats (FFVship.cpp:489-520,
// doc/FFVship.md:115-153) and console statistics (FFVship.cpp:147-200).
//
// Reader threads read frame pairs into pinned buffers taken from a pool
// (Vship_PinnedMalloc) and queue them; one worker thread per GPU handler
// (-g) computes Vship_ComputeSSIMU2 and returns the buffers; an aggregator
// stores the scores. Decoding is replaced by reading raw frames (FFMS2 is not
// available): --width/--height/--pix-fmt describe both files unless the
// --encoded-* variants differ. Only the SSIMULACRA2 metric is provided.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <optional>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/VshipAPI.h"

namespace {

struct PixFmt {
    const char* name;
    int subw, subh, bits;
};

const PixFmt kFormats[] = {
    {"yuv420p", 1, 1, 8},      {"yuv422p", 1, 0, 8},      {"yuv444p", 0, 0, 8},
    {"yuv420p10le", 1, 1, 10}, {"yuv422p10le", 1, 0, 10}, {"yuv444p10le", 0, 0, 10},
    {"yuv420p12le", 1, 1, 12}, {"yuv444p12le", 0, 0, 12}, {"yuv420p16le", 1, 1, 16},
};

struct Video {
    std::string path;
    int width = 0, height = 0;
    const PixFmt* fmt = nullptr;
    int64_t plane_bytes[3] = {0, 0, 0};
    int64_t line[3] = {0, 0, 0};
    int64_t frame_bytes = 0;
    int64_t frames = 0;

    void open(const std::string& p, int w, int h, const PixFmt* f)
    {
        path = p;
        width = w;
        height = h;
        fmt = f;
        const int bpc = f->bits > 8 ? 2 : 1;
        for (int i = 0; i < 3; i++) {
            const int pw = i ? ((w - 1) >> f->subw) + 1 : w;
            const int ph = i ? ((h - 1) >> f->subh) + 1 : h;
            line[i] = (int64_t)pw * bpc;
            plane_bytes[i] = line[i] * ph;
        }
        frame_bytes = plane_bytes[0] + plane_bytes[1] + plane_bytes[2];
        std::ifstream in(path, std::ios::binary | std::ios::ate);
        if (!in) throw std::runtime_error("cannot open " + path);
        frames = (int64_t)in.tellg() / frame_bytes;
    }

    void read(int64_t index, uint8_t* dst, std::ifstream& in) const
    {
        in.seekg(index * frame_bytes);
        in.read(reinterpret_cast<char*>(dst), frame_bytes);
        if (!in) throw std::runtime_error("short read in " + path);
    }
};

// Bounded blocking queue with close().
template <typename T>
class Queue {
public:
    explicit Queue(size_t cap = 0) : cap_(cap) {}
    void push(T v)
    {
        std::unique_lock<std::mutex> lk(m_);
        not_full_.wait(lk, [&] { return cap_ == 0 || q_.size() < cap_ || closed_; });
        q_.push_back(std::move(v));
        not_empty_.notify_one();
    }
    std::optional<T> pop()
    {
        std::unique_lock<std::mutex> lk(m_);
        not_empty_.wait(lk, [&] { return !q_.empty() || closed_; });
        if (q_.empty()) return std::nullopt;
        T v = std::move(q_.front());
        q_.pop_front();
        not_full_.notify_one();
        return v;
    }
    void close()
    {
        std::lock_guard<std::mutex> lk(m_);
        closed_ = true;
        not_empty_.notify_all();
        not_full_.notify_all();
    }

private:
    size_t cap_;
    std::deque<T> q_;
    std::mutex m_;
    std::condition_variable not_empty_, not_full_;
    bool closed_ = false;
};

struct Args {
    std::string source, encoded, json;
    int width = 0, height = 0, ewidth = 0, eheight = 0;
    std::string fmt = "yuv420p", efmt;
    int start = 0, end = -1, every = 1, offset = 0;
    std::vector<int> source_indices, encoded_indices;
    int threads = 1, gpu_threads = 3, gpu_id = 0;
    bool list_gpu = false, live = false, dry_run = false, version = false;
    int crop_src[4] = {0, 0, 0, 0};  // top, bottom, left, right
    int crop_enc[4] = {0, 0, 0, 0};
    int matrix = Vship_MATRIX_BT709, transfer = Vship_TRC_BT709, primaries = Vship_PRIMARIES_BT709;
    int range = Vship_RangeLimited, chroma_loc = Vship_ChromaLoc_Left;
};

void usage()
{
    std::cout <<
        "usage: ffvship_amd --source FILE --encoded FILE --width W --height H [--pix-fmt FMT]\n"
        "                   [--encoded-width W --encoded-height H --encoded-pix-fmt FMT]\n"
        "                   [-m SSIMULACRA2] [--start S] [--end E] [-e|--every N] [--encoded-offset K]\n"
        "                   [--source-indices i,j,..] [--encoded-indices i,j,..]\n"
        "                   [-t THREADS] [-g GPU_THREADS] [--gpu-id ID] [--json OUT] [--list-gpu]\n"
        "                   [--live-score-output] [--crop{Top,Bottom,Left,Right}{Source,Encoded} N]\n"
        "                   [--matrix 709|601|2020] [--transfer 709|srgb|pq|hlg|linear] [--range limited|full]\n"
        "                   [--dry-run] [--version]\n"
        "raw planar YUV input; pix-fmt: yuv420p yuv422p yuv444p yuv420p10le yuv422p10le yuv444p10le\n"
        "yuv420p12le yuv444p12le yuv420p16le\n";
}

std::vector<int> parse_list(const std::string& s)
{
    std::vector<int> v;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ',')) v.push_back(std::stoi(tok));
    return v;
}

Args parse(int argc, char** argv)
{
    Args a;
    auto need = [&](int& i) -> std::string {
        if (i + 1 >= argc) throw std::runtime_error(std::string("missing value after ") + argv[i]);
        return argv[++i];
    };
    for (int i = 1; i < argc; i++) {
        const std::string k = argv[i];
        if (k == "-h" || k == "--help") { usage(); std::exit(0); }
        else if (k == "--source" || k == "--reference") a.source = need(i);
        else if (k == "--encoded" || k == "--distorted") a.encoded = need(i);
        else if (k == "--width") a.width = std::stoi(need(i));
        else if (k == "--height") a.height = std::stoi(need(i));
        else if (k == "--pix-fmt") a.fmt = need(i);
        else if (k == "--encoded-width") a.ewidth = std::stoi(need(i));
        else if (k == "--encoded-height") a.eheight = std::stoi(need(i));
        else if (k == "--encoded-pix-fmt") a.efmt = need(i);
        else if (k == "-m") { const std::string m = need(i); if (m != "SSIMULACRA2") throw std::runtime_error("only -m SSIMULACRA2 is provided"); }
        else if (k == "--start") a.start = std::stoi(need(i));
        else if (k == "--end") a.end = std::stoi(need(i));
        else if (k == "-e" || k == "--every") a.every = std::stoi(need(i));
        else if (k == "--encoded-offset") a.offset = std::stoi(need(i));
        else if (k == "--source-indices") a.source_indices = parse_list(need(i));
        else if (k == "--encoded-indices") a.encoded_indices = parse_list(need(i));
        else if (k == "-t") a.threads = std::stoi(need(i));
        else if (k == "-g") a.gpu_threads = std::stoi(need(i));
        else if (k == "--gpu-id") a.gpu_id = std::stoi(need(i));
        else if (k == "--json") a.json = need(i);
        else if (k == "--list-gpu") a.list_gpu = true;
        else if (k == "--live-score-output") a.live = true;
        else if (k == "--dry-run") a.dry_run = true;
        else if (k == "--version") a.version = true;
        else if (k == "--matrix") { const std::string m = need(i); a.matrix = m == "601" ? Vship_MATRIX_ST170_M : (m == "2020" ? Vship_MATRIX_BT2020_NCL : Vship_MATRIX_BT709); }
        else if (k == "--transfer") {
            const std::string t = need(i);
            a.transfer = t == "srgb" ? Vship_TRC_sRGB : t == "pq" ? Vship_TRC_PQ : t == "hlg" ? Vship_TRC_HLG
                       : t == "linear" ? Vship_TRC_Linear : Vship_TRC_BT709;
        }
        else if (k == "--range") a.range = need(i) == "full" ? Vship_RangeFull : Vship_RangeLimited;
        else if (k.rfind("--crop", 0) == 0) {
            const char* sides[4] = {"Top", "Bottom", "Left", "Right"};
            bool ok = false;
            for (int s = 0; s < 4; s++) {
                if (k == std::string("--crop") + sides[s] + "Source") { a.crop_src[s] = std::stoi(need(i)); ok = true; }
                if (k == std::string("--crop") + sides[s] + "Encoded") { a.crop_enc[s] = std::stoi(need(i)); ok = true; }
            }
            if (!ok) throw std::runtime_error("unknown option " + k);
        }
        else throw std::runtime_error("unknown option " + k);
    }
    return a;
}

const PixFmt* find_format(const std::string& name)
{
    for (const auto& f : kFormats)
        if (name == f.name) return &f;
    throw std::runtime_error("unsupported pix-fmt " + name);
}

Vship_Sample_t sample_of(int bits)
{
    switch (bits) {
    case 8: return Vship_SampleUINT8;
    case 10: return Vship_SampleUINT10;
    case 12: return Vship_SampleUINT12;
    default: return Vship_SampleUINT16;
    }
}

Vship_Colorspace_t colorspace(const Args& a, const Video& v, const int crop[4], int target_w, int target_h)
{
    Vship_Colorspace_t cs;
    std::memset(&cs, 0, sizeof(cs));
    cs.width = v.width;
    cs.height = v.height;
    cs.target_width = target_w == v.width ? -1 : target_w;
    cs.target_height = target_h == v.height ? -1 : target_h;
    cs.sample = sample_of(v.fmt->bits);
    cs.range = (Vship_Range_t)a.range;
    cs.subsampling.subw = v.fmt->subw;
    cs.subsampling.subh = v.fmt->subh;
    cs.chromaLocation = (Vship_ChromaLocation_t)a.chroma_loc;
    cs.colorFamily = Vship_ColorYUV;
    cs.YUVMatrix = (Vship_YUVMatrix_t)a.matrix;
    cs.transferFunction = (Vship_TransferFunction_t)a.transfer;
    cs.primaries = (Vship_Primaries_t)a.primaries;
    cs.crop.top = crop[0];
    cs.crop.bottom = crop[1];
    cs.crop.left = crop[2];
    cs.crop.right = crop[3];
    return cs;
}

// Frame pairs to compare, doc/FFVship.md "start / end / every",
// "encoded-offset" and "source-indices / encoded-indices".
void select_frames(const Args& a, int64_t src_frames, int64_t enc_frames, std::vector<int>& src, std::vector<int>& enc)
{
    if (!a.encoded_indices.empty() && a.source_indices.empty())
        throw std::runtime_error("--encoded-indices needs --source-indices");
    if (a.every < 1) throw std::runtime_error("--every must be >= 1");
    if (!a.source_indices.empty()) {
        const std::vector<int>& ei = a.encoded_indices.empty() ? a.source_indices : a.encoded_indices;
        if (ei.size() != a.source_indices.size())
            throw std::runtime_error("--source-indices and --encoded-indices differ in length");
        for (size_t i = 0; i < ei.size(); i++) {
            const int64_t s = a.start + (int64_t)a.every * a.source_indices[i];
            const int64_t e = a.start + (int64_t)a.every * ei[i] + a.offset;
            if (s < 0 || s >= src_frames || e < 0 || e >= enc_frames) continue;
            src.push_back((int)s);
            enc.push_back((int)e);
        }
        return;
    }
    const int64_t end = a.end < 0 ? src_frames : std::min<int64_t>(a.end, src_frames);
    for (int64_t s = a.start; s < end; s += a.every) {
        const int64_t e = s + a.offset;
        if (s < 0 || e < 0 || e >= enc_frames) continue;
        src.push_back((int)s);
        enc.push_back((int)e);
    }
}

void print_stats(const std::vector<double>& data, const std::string& label)
{
    if (data.empty()) return;
    std::vector<double> s = data;
    std::sort(s.begin(), s.end());
    const size_t n = s.size();
    double mean = 0.0;
    for (double v : s) mean += v;
    mean /= n;
    double sq = 0.0;
    for (double v : s) sq += (v - mean) * (v - mean);
    const std::pair<const char*, double> rows[] = {
        {"Average", mean}, {"Standard Deviation", std::sqrt(sq / n)}, {"Median", s[n / 2]},
        {"5th percentile", s[n / 20]}, {"95th percentile", s[19 * n / 20]}, {"Minimum", s.front()},
        {"Maximum", s.back()}};
    std::cout << "----------------" << label << "---------------\n";
    for (const auto& r : rows)
        std::cout << std::left << std::setw(18) << r.first << " : " << std::right << std::setw(12) << std::fixed
                  << std::setprecision(6) << r.second << "\n";
}

std::string error_text()
{
    const int n = Vship_GetDetailedLastError(nullptr, 0);
    std::string s(n > 0 ? n : 1, '\0');
    Vship_GetDetailedLastError(&s[0], n);
    if (!s.empty() && s.back() == '\0') s.pop_back();
    return s;
}

}  // namespace

int main(int argc, char** argv)
{
    // HIP hardware queues of this process, read when the runtime starts:
    // HIP's default of 4 makes the handler streams share queues (measured on
    // MI355X with 8 handlers: 29.2k pairs/s with 4 queues, 31.7k with 16).
    // A value already in the environment wins.
    setenv("GPU_MAX_HW_QUEUES", "16", 0);
    Args a;
    try {
        a = parse(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        usage();
        return 2;
    }
    if (a.version) {
        const Vship_Version v = Vship_GetVersion();
        std::cout << "ffvship_amd (libvship_amd " << v.major << "." << v.minor << "." << v.minorMinor << ")\n";
        return 0;
    }
    if (a.list_gpu) {
        int n = 0;
        if (Vship_GetDeviceCount(&n) != Vship_NoError) { std::cerr << error_text() << "\n"; return 1; }
        for (int i = 0; i < n; i++) {
            Vship_DeviceInfo info;
            if (Vship_GetDeviceInfo(&info, i) == Vship_NoError) std::cout << "GPU " << i << ": " << info.name << "\n";
        }
        return 0;
    }
    if (a.source.empty() || a.encoded.empty() || a.width <= 0 || a.height <= 0) {
        std::cerr << "error: --source, --encoded, --width and --height are required\n";
        usage();
        return 2;
    }
    Video v1, v2;
    std::vector<int> fsrc, fenc;
    try {
        v1.open(a.source, a.width, a.height, find_format(a.fmt));
        v2.open(a.encoded, a.ewidth ? a.ewidth : a.width, a.eheight ? a.eheight : a.height,
                find_format(a.efmt.empty() ? a.fmt : a.efmt));
        select_frames(a, v1.frames, v2.frames, fsrc, fenc);
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 2;
    }
    const int num_frames = (int)fsrc.size();
    if (a.dry_run) {
        std::cout << num_frames << "\n";
        for (int i = 0; i < num_frames; i++) std::cout << fsrc[i] << " " << fenc[i] << "\n";
        return 0;
    }
    if (num_frames == 0) {
        std::cerr << "error: no frame pair selected\n";
        return 2;
    }
    // the encoded video is resized so that, once cropped, it matches the
    // cropped source (doc/FFVship.md "--crop...")
    const int fw = v1.width - a.crop_src[2] - a.crop_src[3];
    const int fh = v1.height - a.crop_src[0] - a.crop_src[1];
    const int etw = fw + a.crop_enc[2] + a.crop_enc[3];
    const int eth = fh + a.crop_enc[0] + a.crop_enc[1];
    const Vship_Colorspace_t cs1 = colorspace(a, v1, a.crop_src, v1.width, v1.height);
    const Vship_Colorspace_t cs2 = colorspace(a, v2, a.crop_enc, etw, eth);

    if (Vship_SetDevice(a.gpu_id) != Vship_NoError) { std::cerr << error_text() << "\n"; return 1; }
    const int workers = std::max(1, a.gpu_threads);
    std::vector<Vship_SSIMU2Handler> handlers(workers);
    for (int i = 0; i < workers; i++) {
        if (Vship_SSIMU2Init2(&handlers[i], cs1, cs2, a.gpu_id) != Vship_NoError) {
            std::cerr << "error: " << error_text() << "\n";
            return 1;
        }
    }
    // pinned buffer pool: two frames per queued pair plus one pair per worker
    const int readers = std::max(1, a.threads);
    const int pool_pairs = workers + 2 * readers;
    Queue<uint8_t*> pool;
    std::vector<void*> pinned;
    const int64_t buf_bytes = std::max(v1.frame_bytes, v2.frame_bytes);
    for (int i = 0; i < 2 * pool_pairs; i++) {
        void* p = nullptr;
        if (Vship_PinnedMalloc(&p, (uint64_t)buf_bytes) != Vship_NoError) {
            std::cerr << "error: " << error_text() << "\n";
            return 1;
        }
        pinned.push_back(p);
        pool.push(static_cast<uint8_t*>(p));
    }
    struct Task {
        int index;
        uint8_t* src;
        uint8_t* enc;
    };
    Queue<Task> tasks(pool_pairs);
    Queue<std::pair<int, double>> results;
    std::vector<double> scores(num_frames, 0.0);
    int error = 0;
    std::string error_msg;
    std::mutex error_lock;
    auto fail = [&](const std::string& m) {
        std::lock_guard<std::mutex> lk(error_lock);
        if (!error) error_msg = m;
        error = 1;
    };

    if (a.live) std::cout << num_frames << std::endl;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> reader_threads, worker_threads;
    for (int r = 0; r < readers; r++) {
        reader_threads.emplace_back([&, r] {
            std::ifstream in1(v1.path, std::ios::binary), in2(v2.path, std::ios::binary);
            for (int i = num_frames * r / readers; i < num_frames * (r + 1) / readers && !error; i++) {
                uint8_t* s = *pool.pop();
                uint8_t* e = *pool.pop();
                try {
                    v1.read(fsrc[i], s, in1);
                    v2.read(fenc[i], e, in2);
                } catch (const std::exception& ex) {
                    fail(ex.what());
                    pool.push(s);
                    pool.push(e);
                    break;
                }
                tasks.push({i, s, e});
            }
        });
    }
    for (int w = 0; w < workers; w++) {
        worker_threads.emplace_back([&, w] {
            while (true) {
                std::optional<Task> t = tasks.pop();
                if (!t) break;
                if (!error) {
                    const uint8_t* p1[3] = {t->src, t->src + v1.plane_bytes[0], t->src + v1.plane_bytes[0] + v1.plane_bytes[1]};
                    const uint8_t* p2[3] = {t->enc, t->enc + v2.plane_bytes[0], t->enc + v2.plane_bytes[0] + v2.plane_bytes[1]};
                    double s = 0.0;
                    if (Vship_ComputeSSIMU2(handlers[w], &s, p1, p2, v1.line, v2.line) != Vship_NoError) {
                        char buf[1024];
                        Vship_SSIMU2GetDetailedLastError(handlers[w], buf, sizeof(buf));
                        fail(buf);
                    } else {
                        results.push({t->index, s});
                    }
                }
                pool.push(t->src);
                pool.push(t->enc);
            }
        });
    }
    std::thread aggregator([&] {
        while (true) {
            std::optional<std::pair<int, double>> r = results.pop();
            if (!r) break;
            scores[r->first] = r->second;
            if (a.live) std::cout << r->first << " " << std::setprecision(15) << r->second << std::endl;
        }
    });
    for (auto& t : reader_threads) t.join();
    tasks.close();
    for (auto& t : worker_threads) t.join();
    results.close();
    aggregator.join();
    const double seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (auto& h : handlers) Vship_SSIMU2Free(h);
    for (void* p : pinned) Vship_PinnedFree(p);
    if (error) {
        std::cerr << "error: " << error_msg << "\n";
        return 1;
    }

    if (!a.json.empty()) {
        std::ofstream js(a.json);
        if (!js) { std::cerr << "Failed to open output file\n"; return 1; }
        js << std::setprecision(15) << "[";
        for (int i = 0; i < num_frames; i++) js << "[" << scores[i] << (i + 1 < num_frames ? "], " : "]");
        js << "]";
    }
    if (a.live) return 0;
    std::cout << "SSIMU2 Result between " << a.source << " and " << a.encoded << "\n";
    std::cout << "Computed " << num_frames << " frames at " << std::setprecision(6) << num_frames / seconds
              << " fps\n\n";
    print_stats(scores, "SSIMULACRA2");
    return 0;
}
