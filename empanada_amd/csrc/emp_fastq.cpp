#include "emp_fastq.h"

#include <string.h>

#include "../../include/empanada.h"

namespace emp {

namespace {

// Byte classes of a sequence line: 1 = ACGTacgt, 2 = drops the read (a byte in [64,127] with
// bit 3 set, markNs == 9 at preprocess.cu:119-121), 0 = anything else.
struct ByteClass {
    uint8_t c[256];
    ByteClass() {
        for (int i = 0; i < 256; ++i) c[i] = (i >= 64 && i < 128 && (i & 8)) ? 2 : 0;
        const char *ok = "ACGTacgt";
        for (const char *p = ok; *p; ++p) c[(uint8_t)*p] = 1;
    }
};
const ByteClass kClass;

}  // namespace

int parse_fastq_block(const char *data, size_t size, bool final, size_t *consumed, read_fn fn, void *user) {
    *consumed = 0;
    if (size == 0) return EMP_OK;
    if (data[0] != '@') return EMP_ERR_FORMAT;  // preprocess.cu:207
    size_t use = size;
    // parseLines (preprocess.cu:26-37): lines = newlines before the last byte + 1.
    if (!final) {
        uint64_t nl = 0;
        const char *p = data, *end = data + size - 1;
        while (p < end && (p = (const char *)memchr(p, '\n', (size_t)(end - p))) != nullptr) {
            ++nl;
            ++p;
        }
        uint64_t n_lines = nl + 1;
        if (n_lines % 4 != 0 || data[size - 1] != '\n') {
            // resizeBuffer (preprocess.cu:46-62): keep (n_lines-1)/4 records.
            uint64_t keep = ((n_lines - 1) / 4) * 4;
            if (keep == 0) return EMP_ERR_FORMAT;  // a record longer than the block
            uint64_t seen = 0;
            const char *q = data;
            while (true) {
                q = (const char *)memchr(q, '\n', (size_t)(data + size - q));
                if (++seen == keep) break;
                ++q;
            }
            use = (size_t)(q - data) + 1;
        }
    }
    // extractReads + removeLQReads (preprocess.cu:70-182).
    uint64_t line = 0;
    const char *p = data, *end = data + use;
    while (p < end) {
        const char *nlp = (const char *)memchr(p, '\n', (size_t)(end - p));
        const char *le = nlp ? nlp : end;
        if ((line & 3) == 1 && le > p) {
            int cls = 1;
            for (const char *c = p; c < le; ++c) {
                uint8_t k = kClass.c[(uint8_t)*c];
                if (k == 2) { cls = 2; break; }
                if (k == 0) cls = 0;
            }
            if (cls == 0) return EMP_ERR_BASE;
            if (cls == 1) {
                int rc = fn(user, p, (size_t)(le - p));
                if (rc) return rc;
            }
        }
        ++line;
        if (!nlp) break;
        p = nlp + 1;
    }
    *consumed = use;
    return EMP_OK;
}

}  // namespace emp
