// output.bin writer: records in write_edges_binary's layout (overlap.cu:148-161), appended in
// order while later rounds run.
#pragma once

#include <stdint.h>

#include <array>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace emp {

// Per-context resources of direct writes, kept from one emp_run_to_file to the next: a Linux AIO
// context (its teardown waits for an RCU grace period, tens of milliseconds, so it is not paid per
// file) and aligned bounce buffers.
struct DirectIO {
    unsigned long ctx = 0;  // aio_context_t
    bool ready = false;
    bool broken = false;  // its completions could not be read: not used again
    std::vector<uint8_t *> bufs;
    ~DirectIO();
    bool init();
};

// The file is written from two sides at once. Buffered writes into the page cache take the inode's
// lock, so one thread does them all; they reach about 12-14 GB/s on one core, less while the
// rounds run. While they lag behind the records by more than BACKLOG bytes, a fixed share of the
// file's 2 MB chunks among those they have not reached (DIRECT_SET of every 16, while the direct
// side holds fewer than DIRECT_LAG of them) is written with O_DIRECT through Linux AIO instead: a helper thread copies each such chunk into an aligned
// bounce buffer and submits it; the submission returns before the device has the data, so the
// inode lock is not held while it moves. Chunks are page-aligned and each is written by one side
// only, so the page cache never holds a page that a direct write also covers (a chunk whose direct
// write fails is written again, buffered, with the same bytes).
//
// Direct writes need every byte below the chunk's end inside the file's size (no extending
// writes): the blocks ahead of the writes are reserved (fallocate) in OUT_RESERVE steps, as for
// the buffered writes, and cut back to the bytes written when the file is closed. Without
// reservations (file systems other than ext4 and xfs), without O_DIRECT, or on a file that
// existed before (written over in place, then cut to size), every byte is written buffered.
class OutFile {
public:
    static constexpr uint64_t OUT_RESERVE = 64ull << 20;
    static constexpr uint64_t CHUNK = 2ull << 20;
    static constexpr uint32_t DIRECT_SET = 0x2492u;  // chunks c with bit (c % 16) set: 1, 4, 7, 10, 13
    static constexpr uint64_t BACKLOG = 16ull << 20;
    static constexpr uint64_t DIRECT_LAG = 6;  // direct chunks taken and not yet written, at most

    // Opens path as fopen "wb" does. direct: the context's resources (null: buffered only).
    int open(const char *path, DirectIO *direct);
    // Appends one record (n == 0: only its 12-byte head), to be written by flush() or by the
    // direct side. from / to stay valid and unchanged until settle() or finish() returns.
    int append(uint64_t n, int32_t ovlp, const uint32_t *from, const uint32_t *to);
    // Buffered writes of every appended byte not given to the direct side. Appending several
    // records before one flush lets the direct side take a share of them (see BACKLOG).
    int flush();
    // Writes every byte appended so far and waits for the direct writes in flight: afterwards no
    // pointer passed to append() is read again (the caller may free or reuse that memory).
    int settle();
    // Writes what remains, waits for the direct writes, cuts the reservation back and closes.
    int finish();
    ~OutFile();

    uint64_t bytes() const { return avail_; }
    uint64_t direct_bytes() const { return direct_bytes_; }

private:
    struct Piece {
        uint64_t off;
        const uint8_t *p;
        uint64_t len;
    };
    bool designated(uint64_t c) const { return direct_on_ && ((DIRECT_SET >> (c % 16)) & 1u); }
    bool queued(uint64_t c) const { return c < queued_.size() && queued_[c]; }
    int write_buffered(uint64_t limit);
    int drain();
    int gather_write(uint64_t lo, uint64_t hi);
    int pwrite_pieces(const std::vector<Piece> &ps, uint64_t off);
    void pieces_of(uint64_t lo, uint64_t hi, std::vector<Piece> &out);
    void helper();
    int stop_helper();

    int fd_ = -1, dfd_ = -1;
    bool prealloc_ = false, tried_ = false, direct_on_ = false, inplace_ = false;
    uint64_t reserved_ = 0, avail_ = 0, wb_ = 0, next_chunk_ = 0, direct_bytes_ = 0;
    std::deque<std::array<uint8_t, 12>> heads_;
    std::vector<uint8_t> queued_;  // per chunk: 1 when handed to the direct side
    std::vector<Piece> pieces_;  // in file order; guarded by mu_ (the helper reads it)
    DirectIO *dio_ = nullptr;
    // the helper's queue of complete designated chunks
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<uint64_t> queue_;
    bool done_ = false;
    uint64_t taken_ = 0, written_ = 0;  // chunks the helper has taken / whose direct write has completed
    int helper_err_ = 0;
    std::thread helper_th_;
};

}  // namespace emp
