// Included ahead of emp_outfile.cpp (-include) in the failure-injection build of the writer check:
// every fifth AIO submission fails, and one completion in seven reports an I/O error. The writer
// must then write those chunks again through its buffered path.
#pragma once
#include <errno.h>
#include <linux/aio_abi.h>
#include <stdarg.h>
#include <sys/syscall.h>
#include <unistd.h>

#include <atomic>

inline std::atomic<long> fi_submits{0}, fi_failed_submits{0}, fi_events{0}, fi_failed_events{0};

inline long fi_syscall(long nr, ...) {
    va_list ap;
    va_start(ap, nr);
    const long a = va_arg(ap, long), b = va_arg(ap, long), c = va_arg(ap, long), d = va_arg(ap, long),
               e = va_arg(ap, long);
    va_end(ap);
    if (nr == SYS_io_submit && fi_submits++ % 5 == 2) {
        ++fi_failed_submits;
        errno = EAGAIN;
        return -1;
    }
    const long r = syscall(nr, a, b, c, d, e);
    if (nr == SYS_io_getevents && r > 0) {
        struct io_event *ev = reinterpret_cast<struct io_event *>(d);
        for (long i = 0; i < r; ++i)
            if (fi_events++ % 7 == 3) {
                ++fi_failed_events;
                ev[i].res = -EIO;
            }
    }
    return r;
}

#define syscall fi_syscall
