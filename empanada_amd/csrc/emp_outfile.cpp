// output.bin writer (emp_outfile.h): buffered writes from one thread, a share of the file's chunks
// written direct through Linux AIO beside them.
#include "emp_outfile.h"

#include <errno.h>
#include <fcntl.h>
#include <linux/aio_abi.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/syscall.h>
#include <sys/uio.h>
#include <sys/vfs.h>
#include <time.h>
#include <unistd.h>

#include <algorithm>

#include "../../include/empanada.h"

namespace emp {

namespace {
constexpr int NBUF = 8;  // bounce buffers (and direct writes in flight)
constexpr long EXT4_MAGIC = 0xEF53, XFS_MAGIC = 0x58465342;
}  // namespace

DirectIO::~DirectIO() {
    if (ctx) (void)syscall(SYS_io_destroy, (aio_context_t)ctx);
    for (uint8_t *b : bufs) free(b);
}

bool DirectIO::init() {
    if (broken) return false;
    if (ready) return true;
    if (!ctx) {
        aio_context_t c = 0;
        if (syscall(SYS_io_setup, NBUF, &c) != 0) return false;
        ctx = (unsigned long)c;
    }
    while (bufs.size() < (size_t)NBUF) {
        void *p = nullptr;
        if (posix_memalign(&p, 4096, OutFile::CHUNK) != 0) return false;
        bufs.push_back((uint8_t *)p);
    }
    ready = true;
    return true;
}

int OutFile::open(const char *path, DirectIO *direct) {
    struct stat st;
    const bool existed = stat(path, &st) == 0 && st.st_size > 0;
    // fopen "wb". A regular file that is there already is written over in place and cut to the
    // new size when closed, instead of being truncated first: ext4 writes a file truncated to zero
    // and rewritten back to disk when it is closed (auto_da_alloc), which made such a step a third
    // longer. The file left is the one fopen "wb" leaves (the same inode, bytes and size).
    inplace_ = existed && S_ISREG(st.st_mode);
    fd_ = ::open(path, O_WRONLY | O_CREAT | (inplace_ ? 0 : O_TRUNC) | O_CLOEXEC, 0666);
    if (fd_ < 0) return EMP_ERR_IO;
    struct statfs sf;
    prealloc_ = fstatfs(fd_, &sf) == 0 && ((long)sf.f_type == EXT4_MAGIC || (long)sf.f_type == XFS_MAGIC);
    const char *env = getenv("EMPANADA_DIRECT_WRITES");
    const bool allowed = !(env && env[0] == '0');
    if (prealloc_ && !existed && direct && allowed && direct->init()) {
        dfd_ = ::open(path, O_WRONLY | O_DIRECT | O_CLOEXEC);
        if (dfd_ >= 0) {
            dio_ = direct;
            direct_on_ = true;
            helper_th_ = std::thread([this]() { helper(); });
        }
    }
    return EMP_OK;
}

OutFile::~OutFile() {
    if (helper_th_.joinable()) (void)stop_helper();
    if (dfd_ >= 0) close(dfd_);
    if (fd_ >= 0) close(fd_);
}

int OutFile::append(uint64_t n, int32_t ovlp, const uint32_t *from, const uint32_t *to) {
    const uint64_t bytes = 12 + 8 * n;
    if (prealloc_ && avail_ + bytes > reserved_) {
        const uint64_t end = std::max(avail_ + bytes, reserved_ + OUT_RESERVE);
        tried_ = true;
        if (fallocate(fd_, 0, (off_t)reserved_, (off_t)(end - reserved_)) == 0) {
            reserved_ = end;
        } else {
            // the writes go on without reservations, and so without direct writes (those
            // below reserved_ are already queued and stay valid)
            prealloc_ = false;
            direct_on_ = false;
        }
    }
    std::array<uint8_t, 12> &h = heads_.emplace_back();
    memcpy(h.data(), &n, 8);
    memcpy(h.data() + 8, &ovlp, 4);
    {
        std::lock_guard<std::mutex> g(mu_);
        pieces_.push_back(Piece{avail_, h.data(), 12});
        if (n) {
            pieces_.push_back(Piece{avail_ + 12, reinterpret_cast<const uint8_t *>(from), 4 * n});
            pieces_.push_back(Piece{avail_ + 12 + 4 * n, reinterpret_cast<const uint8_t *>(to), 4 * n});
        }
    }
    avail_ += bytes;
    // While the buffered writes lag behind by more than BACKLOG, the designated chunks among the
    // complete ones they have not reached go to the direct side. Otherwise every byte is written
    // buffered: a direct write takes about a millisecond to complete, which would only add to the
    // end of the file once the buffered writes keep up.
    // The direct side never holds more than DIRECT_LAG chunks not yet written: when the device
    // slows down (a disk's cache full after gigabytes of direct writes), later designated chunks
    // go buffered instead of waiting in its queue, and the end of the file never waits for more
    // than DIRECT_LAG direct chunks.
    next_chunk_ = std::max(next_chunk_, (wb_ + CHUNK - 1) / CHUNK);
    const bool behind = avail_ - wb_ > BACKLOG;
    for (; (next_chunk_ + 1) * CHUNK <= avail_; ++next_chunk_) {
        if (!behind || !designated(next_chunk_)) continue;
        std::lock_guard<std::mutex> g(mu_);
        if (taken_ - written_ >= DIRECT_LAG) continue;
        if (queued_.size() <= next_chunk_) queued_.resize(next_chunk_ + 1, 0);
        queued_[next_chunk_] = 1;
        direct_bytes_ += CHUNK;
        queue_.push_back(next_chunk_);
        ++taken_;
        cv_.notify_all();
    }
    return EMP_OK;
}

int OutFile::flush() { return write_buffered(avail_); }

// Buffered writes of [wb_, limit) minus the chunks of the direct side.
int OutFile::write_buffered(uint64_t limit) {
    while (wb_ < limit) {
        const uint64_t c = wb_ / CHUNK;
        if (queued(c)) {
            wb_ = (c + 1) * CHUNK;
            continue;
        }
        uint64_t end = limit;
        for (uint64_t k = c + 1; k * CHUNK < limit; ++k) {
            if (queued(k)) {
                end = k * CHUNK;
                break;
            }
        }
        const int rc = gather_write(wb_, end);
        if (rc) return rc;
        wb_ = end;
    }
    return EMP_OK;
}

// The pieces that cover file bytes [lo, hi), clipped to it.
void OutFile::pieces_of(uint64_t lo, uint64_t hi, std::vector<Piece> &out) {
    out.clear();
    size_t a = 0, b = pieces_.size();
    while (a < b) {  // first piece ending past lo
        const size_t m = (a + b) / 2;
        if (pieces_[m].off + pieces_[m].len <= lo) a = m + 1;
        else b = m;
    }
    for (size_t i = a; i < pieces_.size() && pieces_[i].off < hi; ++i) {
        const Piece &p = pieces_[i];
        const uint64_t s = std::max(lo, p.off), e = std::min(hi, p.off + p.len);
        if (e > s) out.push_back(Piece{s, p.p + (s - p.off), e - s});
    }
}

int OutFile::gather_write(uint64_t lo, uint64_t hi) {
    std::vector<Piece> ps;
    pieces_of(lo, hi, ps);  // the writer thread is the only one that changes pieces_
    return pwrite_pieces(ps, lo);
}

// The pieces' bytes, back to back, at file offset off (buffered).
int OutFile::pwrite_pieces(const std::vector<Piece> &ps, uint64_t off) {
    constexpr size_t MAXV = 1024;
    struct iovec iov[MAXV];
    size_t i = 0;
    while (i < ps.size()) {
        int cnt = 0;
        for (; i < ps.size() && cnt < (int)MAXV; ++i, ++cnt) iov[cnt] = iovec{(void *)ps[i].p, (size_t)ps[i].len};
        struct iovec *v = iov;
        while (cnt > 0) {
            const ssize_t w = pwritev(fd_, v, cnt, (off_t)off);
            if (w < 0) {
                if (errno == EINTR) continue;
                return EMP_ERR_IO;
            }
            off += (uint64_t)w;
            size_t left = (size_t)w;
            while (cnt > 0 && left >= v->iov_len) {
                left -= v->iov_len;
                ++v;
                --cnt;
            }
            if (cnt > 0) {
                v->iov_base = (char *)v->iov_base + left;
                v->iov_len -= left;
            }
        }
    }
    return EMP_OK;
}

// Direct side: copies each queued chunk into a free bounce buffer and submits it; reaps
// completions while it waits for buffers or for more chunks. A chunk whose direct write fails (at
// submission or at completion) is written again, buffered, from its bounce buffer (still intact);
// if the completions cannot be read at all, every chunk in flight is, and from then on every chunk
// the helper takes is written buffered from the records. Only a failed buffered write fails the
// file.
void OutFile::helper() {
    const aio_context_t ctx = (aio_context_t)dio_->ctx;
    struct io_event ev[NBUF];
    std::vector<int> freeb;
    for (int i = 0; i < NBUF; ++i) freeb.push_back(i);
    uint64_t chunk_of[NBUF] = {};
    bool busy[NBUF] = {};
    int inflight = 0, err = 0;
    bool broken = false;
    std::vector<Piece> ps;
    auto finished = [&](uint64_t n) {
        std::lock_guard<std::mutex> g(mu_);
        written_ += n;
        cv_.notify_all();
    };
    auto rewrite = [&](int bi) {  // the chunk in bounce buffer bi, buffered
        const std::vector<Piece> one{Piece{chunk_of[bi] * CHUNK, dio_->bufs[bi], CHUNK}};
        if (pwrite_pieces(one, chunk_of[bi] * CHUNK) != EMP_OK) err = EMP_ERR_IO;
    };
    auto reap = [&](long min_nr, struct timespec *to) {
        const long r = syscall(SYS_io_getevents, ctx, min_nr, (long)NBUF, ev, to);
        if (r < 0) {
            if (errno == EINTR) return;
            for (int i = 0; i < NBUF; ++i)
                if (busy[i]) rewrite(i);  // the buffers stay out of use: a write may still read them
            finished((uint64_t)inflight);
            inflight = 0;
            broken = true;
            dio_->broken = true;  // it may still hold events of this file's writes
            return;
        }
        for (long i = 0; i < r; ++i) {
            const int bi = (int)ev[i].data;
            if ((long long)ev[i].res != (long long)CHUNK) rewrite(bi);
            busy[bi] = false;
            freeb.push_back(bi);
        }
        inflight -= (int)r;
        if (r > 0) finished((uint64_t)r);
    };
    for (;;) {
        uint64_t c = 0;
        bool have = false, stop = false;
        {
            std::unique_lock<std::mutex> g(mu_);
            if (inflight == 0) cv_.wait(g, [this]() { return !queue_.empty() || done_; });
            if (!queue_.empty()) {
                c = queue_.front();
                queue_.pop_front();
                have = true;
                pieces_of(c * CHUNK, (c + 1) * CHUNK, ps);
            } else if (done_ && inflight == 0) {
                stop = true;
            }
        }
        if (stop) break;
        if (!have) {
            struct timespec to{0, 200000};
            reap(1, &to);
            continue;
        }
        while (!broken && freeb.empty() && inflight > 0) reap(1, nullptr);
        if (broken || freeb.empty()) {
            if (pwrite_pieces(ps, c * CHUNK) != EMP_OK) err = EMP_ERR_IO;
            finished(1);
            continue;
        }
        const int bi = freeb.back();
        freeb.pop_back();
        uint8_t *buf = dio_->bufs[bi];
        for (const Piece &p : ps) memcpy(buf + (p.off - c * CHUNK), p.p, p.len);
        chunk_of[bi] = c;
        struct iocb cb;
        memset(&cb, 0, sizeof cb);
        cb.aio_data = (uint64_t)bi;
        cb.aio_lio_opcode = IOCB_CMD_PWRITE;
        cb.aio_fildes = (uint32_t)dfd_;
        cb.aio_buf = (uint64_t)(uintptr_t)buf;
        cb.aio_nbytes = CHUNK;
        cb.aio_offset = (int64_t)(c * CHUNK);
        struct iocb *pcb = &cb;
        if (syscall(SYS_io_submit, ctx, 1L, &pcb) == 1) {
            busy[bi] = true;
            ++inflight;
        } else {
            rewrite(bi);
            freeb.push_back(bi);
            finished(1);
        }
    }
    std::lock_guard<std::mutex> g(mu_);
    helper_err_ = err;
    cv_.notify_all();
}

int OutFile::drain() {
    if (!helper_th_.joinable()) return EMP_OK;
    std::unique_lock<std::mutex> g(mu_);
    cv_.wait(g, [this]() { return written_ >= taken_; });
    return helper_err_;
}

int OutFile::stop_helper() {
    {
        std::lock_guard<std::mutex> g(mu_);
        done_ = true;
        cv_.notify_all();
    }
    helper_th_.join();
    return helper_err_;
}

int OutFile::settle() {
    const int rc = write_buffered(avail_);
    const int rd = drain();
    return rc ? rc : rd;
}

int OutFile::finish() {
    int rc = write_buffered(avail_);
    if (helper_th_.joinable()) {
        const int rd = stop_helper();
        if (!rc) rc = rd;
    }
    if ((tried_ || inplace_) && ftruncate(fd_, (off_t)avail_) != 0 && !rc) rc = EMP_ERR_IO;
    if (dfd_ >= 0 && close(dfd_) != 0 && !rc) rc = EMP_ERR_IO;
    dfd_ = -1;
    if (fd_ >= 0 && close(fd_) != 0 && !rc) rc = EMP_ERR_IO;
    fd_ = -1;
    return rc;
}

}  // namespace emp
