// On-disk output in the reference's layout (include/io_operators_tmpl.hpp:44-212
// of the reference); see io_operators_tmpl.hpp for the files written.
#include <hip/hip_runtime.h>

#include <charconv>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <fstream>
#include <iostream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>

#include "cfdarcho_main_3d.hpp"
#include "fvm3d.hpp"
#include "io_operators_tmpl.hpp"
#include "mesh3d.hpp"

namespace cfdarco::io {

namespace {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// RFC 4648 base64 of n bytes, appended to out
void base64(const unsigned char* p, size_t n, std::string& out) {
    static const char* A = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    out.reserve(out.size() + 4 * ((n + 2) / 3));
    size_t i = 0;
    for (; i + 3 <= n; i += 3) {
        const unsigned v = (unsigned(p[i]) << 16) | (unsigned(p[i + 1]) << 8) | unsigned(p[i + 2]);
        out += A[v >> 18];
        out += A[(v >> 12) & 63];
        out += A[(v >> 6) & 63];
        out += A[v & 63];
    }
    if (i < n) {
        const unsigned v = (unsigned(p[i]) << 16) | (i + 1 < n ? unsigned(p[i + 1]) << 8 : 0u);
        out += A[v >> 18];
        out += A[(v >> 12) & 63];
        out += i + 1 < n ? A[(v >> 6) & 63] : '=';
        out += '=';
    }
}

std::ofstream open_out(const fs::path& file, std::ios::openmode mode = std::ios::out) {
    std::ofstream f(file, mode);
    if (!f.is_open()) throw std::runtime_error("cannot open " + file.string() + " for writing");
    return f;
}

void put_num(std::string& s, double v) {
    char buf[32];
    const auto r = std::to_chars(buf, buf + sizeof(buf), v);
    s.append(buf, r.ptr);
}
void put_num(std::string& s, size_t v) {
    char buf[24];
    const auto r = std::to_chars(buf, buf + sizeof(buf), v);
    s.append(buf, r.ptr);
}

}  // namespace

void write_vti(const fs::path& file, size_t X, size_t Y, size_t Z, const float* cells) {
    // point (x, y, z) of the image is cell (x, y, z) of the mesh; VTK orders
    // points x fastest, the mesh numbers cells z fastest
    std::vector<float> img(X * Y * Z);
    size_t k = 0;
    for (size_t z = 0; z < Z; ++z)
        for (size_t y = 0; y < Y; ++y)
            for (size_t x = 0; x < X; ++x) img[k++] = cells[Z * Y * x + Z * y + z];
    const uint64_t nbytes = uint64_t(img.size() * sizeof(float));
    std::string s;
    s += "<?xml version=\"1.0\"?>\n";
    s += "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    const std::string ext = "0 " + std::to_string(X - 1) + " 0 " + std::to_string(Y - 1) + " 0 " + std::to_string(Z - 1);
    s += "  <ImageData WholeExtent=\"" + ext + "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n";
    s += "    <Piece Extent=\"" + ext + "\">\n";
    s += "      <PointData Scalars=\"ImageScalars\">\n";
    s += "        <DataArray type=\"Float32\" Name=\"ImageScalars\" format=\"binary\">\n          ";
    // inline binary: the byte-count header and the data, each base64-encoded
    base64(reinterpret_cast<const unsigned char*>(&nbytes), sizeof(nbytes), s);
    base64(reinterpret_cast<const unsigned char*>(img.data()), size_t(nbytes), s);
    s += "\n        </DataArray>\n      </PointData>\n      <CellData>\n      </CellData>\n";
    s += "    </Piece>\n  </ImageData>\n</VTKFile>\n";
    auto f = open_out(file);
    f.write(s.data(), std::streamsize(s.size()));
}

void write_bin(const fs::path& file, const float* values, size_t n) {
    auto f = open_out(file, std::ios::out | std::ios::binary);
    f.write(reinterpret_cast<const char*>(values), std::streamsize(n * sizeof(float)));
    if (!f) throw std::runtime_error("short write to " + file.string());
}

void link_latest(const fs::path& target, const fs::path& link) {
    const fs::path abs_link = fs::absolute(link);
    std::error_code ec;
    fs::remove_all(abs_link, ec);
    fs::create_symlink(fs::absolute(target), abs_link);
}

// ---------------------------------------------------------------------------
// asynchronous stepping snapshots
// ---------------------------------------------------------------------------
namespace {

struct Job {
    float* host;
    size_t n;
    hipEvent_t copied;
    fs::path file;
};

class Stepper {
public:
    Stepper(std::vector<Variable*> vars, fs::path dir) : vars_(std::move(vars)), dir_(std::move(dir)) {
        hip_check(hipStreamCreateWithFlags(&io_, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        hip_check(hipEventCreateWithFlags(&ready_, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&last_, hipEventDisableTiming), "hipEventCreate");
        writer_ = std::thread([this] { write_loop(); });
    }

    ~Stepper() {
        try {
            finish();
        } catch (...) {
        }
    }

    void snapshot(int step) {
        if (error_) rethrow();
        // a step's output buffer is reused as an output two steps later: the
        // previous snapshot's copies finish before anything queued after this
        // point runs on the compute (null) stream
        if (have_last_) hip_check(hipStreamWaitEvent(nullptr, last_, 0), "hipStreamWaitEvent");
        hip_check(hipEventRecord(ready_, nullptr), "hipEventRecord");
        hip_check(hipStreamWaitEvent(io_, ready_, 0), "hipStreamWaitEvent");
        for (Variable* v : vars_) {
            const size_t n = v->current_cu._size;
            Job j{acquire(n), n, nullptr, dir_ / v->name / (std::to_string(step) + ".bin")};
            hip_check(hipEventCreateWithFlags(&j.copied, hipEventDisableTiming), "hipEventCreate");
            if (n)
                hip_check(hipMemcpyAsync(j.host, v->current_cu.data.get(), n * sizeof(float), hipMemcpyDeviceToHost,
                                         io_),
                          "hipMemcpyAsync");
            hip_check(hipEventRecord(j.copied, io_), "hipEventRecord");
            std::lock_guard<std::mutex> lk(mu_);
            jobs_.push_back(std::move(j));
            cv_.notify_all();
        }
        hip_check(hipEventRecord(last_, io_), "hipEventRecord");
        have_last_ = true;
    }

    void finish() {
        if (done_) return;
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            cv_.notify_all();
        }
        if (writer_.joinable()) writer_.join();
        (void)hipStreamSynchronize(io_);
        for (float* p : free_) (void)hipHostFree(p);
        free_.clear();
        (void)hipEventDestroy(ready_);
        (void)hipEventDestroy(last_);
        (void)hipStreamDestroy(io_);
        done_ = true;
        if (error_) rethrow();
    }

private:
    // a page-locked buffer of n floats: a free one of that size, a new one
    // while fewer than kMaxBuffers exist, else the next one the writer frees
    float* acquire(size_t n) {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            for (size_t i = 0; i < free_.size(); ++i)
                if (free_n_[i] == n) {
                    float* p = free_[i];
                    free_.erase(free_.begin() + long(i));
                    free_n_.erase(free_n_.begin() + long(i));
                    return p;
                }
            if (allocated_ < kMaxBuffers || (free_.empty() && jobs_.empty() && !busy_)) {
                ++allocated_;
                lk.unlock();
                void* p = nullptr;
                hip_check(hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(float), hipHostMallocDefault),
                          "hipHostMalloc");
                return static_cast<float*>(p);
            }
            if (!free_.empty()) {  // a buffer of another size: replace it
                float* p = free_.back();
                free_.pop_back();
                free_n_.pop_back();
                lk.unlock();
                (void)hipHostFree(p);
                void* q = nullptr;
                hip_check(hipHostMalloc(&q, std::max<size_t>(n, 1) * sizeof(float), hipHostMallocDefault),
                          "hipHostMalloc");
                return static_cast<float*>(q);
            }
            cv_.wait(lk);
        }
    }

    void write_loop() {
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
                if (jobs_.empty()) return;
                j = std::move(jobs_.front());
                jobs_.pop_front();
                busy_ = true;
            }
            try {
                hip_check(hipEventSynchronize(j.copied), "hipEventSynchronize");
                write_bin(j.file, j.host, j.n);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu_);
                if (!error_) error_ = std::current_exception();
            }
            (void)hipEventDestroy(j.copied);
            std::lock_guard<std::mutex> lk(mu_);
            free_.push_back(j.host);
            free_n_.push_back(j.n);
            busy_ = false;
            cv_.notify_all();
        }
    }

    void rethrow() {
        std::exception_ptr e;
        {
            std::lock_guard<std::mutex> lk(mu_);
            std::swap(e, error_);
        }
        if (e) std::rethrow_exception(e);
    }

    static constexpr int kMaxBuffers = 16;
    std::vector<Variable*> vars_;
    fs::path dir_;
    hipStream_t io_ = nullptr;
    hipEvent_t ready_ = nullptr, last_ = nullptr;
    bool have_last_ = false;
    std::thread writer_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Job> jobs_;
    std::vector<float*> free_;
    std::vector<size_t> free_n_;
    int allocated_ = 0;
    bool busy_ = false;
    bool stop_ = false;
    bool done_ = false;
    std::exception_ptr error_;
};

std::unique_ptr<Stepper> g_stepper;

}  // namespace

bool stepping_active() { return g_stepper != nullptr; }

void stepping_snapshot(int step) {
    if (g_stepper) g_stepper->snapshot(step);
}

void start_stepping(const std::vector<Variable*>& vars, const fs::path& dir) {
    g_stepper = std::make_unique<Stepper>(vars, dir);
}

void stop_stepping() {
    if (!g_stepper) return;
    std::unique_ptr<Stepper> s = std::move(g_stepper);
    s->finish();
}

}  // namespace cfdarco::io

using namespace cfdarco::io;

void store_history(const std::vector<Variable*>& vars_to_store, const Mesh3D* mesh) {
    fs::create_directories(CFDArcoGlobalInit::store_dir);
    for (Variable* var : vars_to_store) {
        const fs::path dir = CFDArcoGlobalInit::store_dir / var->name;
        fs::create_directories(dir);
        // every entry but the last, as the reference's loop bound
        // (i < history.size() - 1) has it
        for (size_t i = 0; i + 1 < var->history.size(); ++i)
            write_vti(dir / ("res_" + std::to_string(i) + ".vti"), mesh->_x, mesh->_y, mesh->_z,
                      var->history[i].data());
    }
    link_latest(CFDArcoGlobalInit::store_dir, CFDArcoGlobalInit::store_dir / ".." / "run_latest");
}

void store_mesh(const Mesh3D* mesh, const fs::path& store_path) {
    // vertexes, faces and nodes in the reference's numbering
    // (src/mesh3d.cpp:323-416): vertex coordinates are x*dx (float) plus the
    // mesh's vertex offsets; a face's quad and its cells (lower cell first) follow
    // from its family and lattice position
    const size_t X = mesh->_x, Y = mesh->_y, Z = mesh->_z;
    const auto V = [&](size_t x, size_t y, size_t z) { return (x * (Y + 1) + y) * (Z + 1) + z; };
    const auto C = [&](size_t x, size_t y, size_t z) { return Z * Y * x + Z * y + z; };
    const std::vector<double>& off = mesh->vertex_offsets();
    std::ofstream f = open_out(store_path / "mesh.json");
    std::string s;
    auto flush = [&] {
        f.write(s.data(), std::streamsize(s.size()));
        s.clear();
    };
    auto ids = [&](const size_t* a, int n) {
        s += '[';
        for (int i = 0; i < n; ++i) {
            if (i) s += ", ";
            put_num(s, a[i]);
        }
        s += ']';
    };
    s += "{\n    \"face\": [\n";
    bool first = true;
    auto face = [&](const size_t q[4], const size_t* cells, int nc) {
        s += first ? "        " : ",\n        ";
        first = false;
        s += "{\"nodes_id\": ";
        ids(cells, nc);
        s += ", \"vertexes_id\": ";
        ids(q, 4);
        s += '}';
        if (s.size() > (1u << 20)) flush();
    };
    for (size_t x = 0; x < X; ++x)  // z-faces: id = (Z+1)*Y*x + (Z+1)*y + zf
        for (size_t y = 0; y < Y; ++y)
            for (size_t zf = 0; zf <= Z; ++zf) {
                const size_t q[4] = {V(x, y, zf), V(x, y + 1, zf), V(x + 1, y + 1, zf), V(x + 1, y, zf)};
                size_t c[2];
                int nc = 0;
                if (zf > 0) c[nc++] = C(x, y, zf - 1);
                if (zf < Z) c[nc++] = C(x, y, zf);
                face(q, c, nc);
            }
    for (size_t z = 0; z < Z; ++z)  // y-faces: id = base + X*(Y+1)*z + (Y+1)*x + yf
        for (size_t x = 0; x < X; ++x)
            for (size_t yf = 0; yf <= Y; ++yf) {
                const size_t q[4] = {V(x, yf, z), V(x, yf, z + 1), V(x + 1, yf, z + 1), V(x + 1, yf, z)};
                size_t c[2];
                int nc = 0;
                if (yf > 0) c[nc++] = C(x, yf - 1, z);
                if (yf < Y) c[nc++] = C(x, yf, z);
                face(q, c, nc);
            }
    for (size_t y = 0; y < Y; ++y)  // x-faces: id = base + (X+1)*Z*y + (X+1)*z + xf
        for (size_t z = 0; z < Z; ++z)
            for (size_t xf = 0; xf <= X; ++xf) {
                const size_t q[4] = {V(xf, y, z + 1), V(xf, y + 1, z + 1), V(xf, y + 1, z), V(xf, y, z)};
                size_t c[2];
                int nc = 0;
                if (xf > 0) c[nc++] = C(xf - 1, y, z);
                if (xf < X) c[nc++] = C(xf, y, z);
                face(q, c, nc);
            }
    s += "\n    ],\n    \"lx\": ";
    put_num(s, double(mesh->_lx));
    s += ",\n    \"ly\": ";
    put_num(s, double(mesh->_ly));
    s += ",\n    \"lz\": ";
    put_num(s, double(mesh->_lz));
    s += ",\n    \"nodes\": [\n";
    first = true;
    for (size_t x = 0; x < X; ++x)
        for (size_t y = 0; y < Y; ++y)
            for (size_t z = 0; z < Z; ++z) {
                s += first ? "        " : ",\n        ";
                first = false;
                const auto fc = mesh->square_node_coord_to_face_idx(x, y, z);
                const auto vx = mesh->square_node_coord_to_vertex_idx(x, y, z);
                s += "{\"face\": ";
                ids(fc.data(), 6);
                s += ", \"vertexes\": ";
                ids(vx.data(), 8);
                s += '}';
                if (s.size() > (1u << 20)) flush();
            }
    s += "\n    ],\n    \"vertexes\": [\n";
    first = true;
    const float d[3] = {mesh->_dx, mesh->_dy, mesh->_dz};
    for (size_t x = 0; x <= X; ++x)
        for (size_t y = 0; y <= Y; ++y)
            for (size_t z = 0; z <= Z; ++z) {
                const size_t v = V(x, y, z);
                const size_t lat[3] = {x, y, z};
                s += first ? "        [" : ",\n        [";
                first = false;
                for (int k = 0; k < 3; ++k) {
                    float p = float(lat[k]) * d[k];
                    if (!off.empty()) p = p + float(off[3 * v + k]);
                    if (k) s += ", ";
                    put_num(s, double(p));
                }
                s += ']';
                if (s.size() > (1u << 20)) flush();
            }
    s += "\n    ],\n    \"x\": ";
    put_num(s, X);
    s += ",\n    \"y\": ";
    put_num(s, Y);
    s += ",\n    \"z\": ";
    put_num(s, Z);
    s += "\n}\n";
    flush();
}

void init_store_history_stepping(const std::vector<Variable*>& vars_to_store, const Mesh3D* mesh,
                                 const fs::path&) {
    CFDArcoGlobalInit::store_stepping = true;
    fs::create_directories(CFDArcoGlobalInit::store_dir);
    for (Variable* v : vars_to_store) fs::create_directories(CFDArcoGlobalInit::store_dir / v->name);
    store_mesh(mesh, CFDArcoGlobalInit::store_dir);
    start_stepping(vars_to_store, CFDArcoGlobalInit::store_dir);
}

void store_history_stepping(const std::vector<Variable*>& vars_to_store, const Mesh3D*, int i) {
    // synchronous form for callers outside the time loop
    for (Variable* v : vars_to_store) {
        const fs::path dir = CFDArcoGlobalInit::store_dir / v->name;
        fs::create_directories(dir);
        Eigen::Matrix<float, -1, 1> h = v->current_cu.to_eigen();
        write_bin(dir / (std::to_string(i) + ".bin"), h.data(), size_t(h.size()));
    }
}

void finalize_history_stepping(const fs::path& store_path) {
    stop_stepping();
    link_latest(CFDArcoGlobalInit::store_dir, store_path / "run_latest");
}
