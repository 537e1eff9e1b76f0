// FASTQ block parsing with the reference's read_block rules (preprocess.cu:195-237).
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace emp {

// Receives each kept read (bytes without the newline). Returns 0 or an EMP_ERR_* code.
typedef int (*read_fn)(void *user, const char *seq, size_t len);

// Parses one block as read_block does: the block must start with '@'; unless `final`, only the
// complete records are used (a record is 4 lines, the last ending in '\n'), *consumed is set to
// their byte count and the caller resumes there. Sequence lines (line % 4 == 1) holding a byte
// in [64,127] with bit 3 set (N, n, K, M, ...) are dropped (preprocess.cu:119-121,150-155);
// empty ones vanish; kept reads must be ACGTacgt (EMP_ERR_BASE otherwise).
int parse_fastq_block(const char *data, size_t size, bool final, size_t *consumed, read_fn fn, void *user);

}  // namespace emp
