// Zip container of GRAL's flow-field files (GFFWrite.cs:171-177, :105-109): one
// entry, named like the file, deflated.  The payload is deflated in parallel
// chunks: each chunk is an independent raw deflate stream that ends on a byte
// boundary with a sync flush (an empty stored block), the last one finishes the
// stream, so their concatenation is one valid deflate stream; the CRC-32 of the
// payload is combined from the chunks' CRCs.  No ZIP64: payload and compressed
// size must stay below 4 GiB (a 1000 x 1000 x 30 field packs to 186 MB).
#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <thread>
#include <vector>

#include "../../include/gral_pf.h"

namespace {

constexpr size_t kChunk = size_t(4) << 20;

bool deflate_chunk(const uint8_t *p, size_t n, bool last, std::vector<uint8_t> &out) {
    z_stream zs;
    std::memset(&zs, 0, sizeof zs);
    if (deflateInit2(&zs, Z_DEFAULT_COMPRESSION, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) return false;
    out.resize(deflateBound(&zs, (uLong)n) + 64);
    zs.next_in = const_cast<Bytef *>(p);
    zs.avail_in = (uInt)n;
    const int flush = last ? Z_FINISH : Z_SYNC_FLUSH;
    for (;;) {
        if (zs.total_out + 64 > out.size()) out.resize(out.size() * 2);
        zs.next_out = out.data() + zs.total_out;
        zs.avail_out = (uInt)(out.size() - zs.total_out);
        const int rc = deflate(&zs, flush);
        if (rc == Z_STREAM_ERROR) { deflateEnd(&zs); return false; }
        if (last ? rc == Z_STREAM_END : (zs.avail_in == 0 && zs.avail_out != 0)) break;
    }
    out.resize(zs.total_out);
    deflateEnd(&zs);
    return true;
}

void put16(std::vector<uint8_t> &b, uint32_t v) { b.push_back(v & 0xff); b.push_back((v >> 8) & 0xff); }
void put32(std::vector<uint8_t> &b, uint32_t v) { put16(b, v & 0xffff); put16(b, v >> 16); }

}  // namespace

extern "C" int gral_pf_zip_write(const char *path, const char *entry, const void *data, size_t len) {
    if (!path || !entry || (!data && len)) return GRAL_PF_EINVAL;
    if (len >= 0xffffffffull) return GRAL_PF_EINVAL;
    const size_t name_len = std::strlen(entry);
    if (name_len == 0 || name_len > 0xffff) return GRAL_PF_EINVAL;
    const uint8_t *src = static_cast<const uint8_t *>(data);
    const size_t nch = std::max<size_t>(1, (len + kChunk - 1) / kChunk);
    std::vector<std::vector<uint8_t>> out(nch);
    std::vector<uLong> crc(nch, 0);
    std::vector<int> ok(nch, 0);
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t nthr = std::min<size_t>(nch, hw);
    auto work = [&](size_t t0) {
        for (size_t c = t0; c < nch; c += nthr) {
            const size_t off = c * kChunk, n = std::min(kChunk, len - std::min(len, off));
            ok[c] = deflate_chunk(src + off, n, c + 1 == nch, out[c]) ? 1 : 0;
            crc[c] = crc32(0L, src + off, (uInt)n);
        }
    };
    std::vector<std::thread> th;
    for (size_t t = 1; t < nthr; ++t) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    uLong total_crc = crc[0];
    size_t csize = out[0].size();
    for (size_t c = 1; c < nch; ++c) {
        const size_t n = std::min(kChunk, len - c * kChunk);
        total_crc = crc32_combine(total_crc, crc[c], (z_off_t)n);
        csize += out[c].size();
    }
    for (size_t c = 0; c < nch; ++c)
        if (!ok[c]) return GRAL_PF_ENOMEM;
    if (csize >= 0xffffffffull) return GRAL_PF_EINVAL;

    std::time_t now = std::time(nullptr);
    std::tm lt;
    localtime_r(&now, &lt);
    const uint32_t dos_time = (uint32_t)((lt.tm_hour << 11) | (lt.tm_min << 5) | (lt.tm_sec / 2));
    const uint32_t dos_date = (uint32_t)(((std::max(lt.tm_year, 80) - 80) << 9) | ((lt.tm_mon + 1) << 5) | lt.tm_mday);

    std::vector<uint8_t> lh;                  // local file header
    put32(lh, 0x04034b50); put16(lh, 20); put16(lh, 0); put16(lh, 8); put16(lh, dos_time); put16(lh, dos_date);
    put32(lh, (uint32_t)total_crc); put32(lh, (uint32_t)csize); put32(lh, (uint32_t)len);
    put16(lh, (uint32_t)name_len); put16(lh, 0);
    lh.insert(lh.end(), entry, entry + name_len);
    std::vector<uint8_t> cd;                  // central directory and its end record
    put32(cd, 0x02014b50); put16(cd, 20); put16(cd, 20); put16(cd, 0); put16(cd, 8); put16(cd, dos_time);
    put16(cd, dos_date); put32(cd, (uint32_t)total_crc); put32(cd, (uint32_t)csize); put32(cd, (uint32_t)len);
    put16(cd, (uint32_t)name_len); put16(cd, 0); put16(cd, 0); put16(cd, 0); put16(cd, 0); put32(cd, 0); put32(cd, 0);
    cd.insert(cd.end(), entry, entry + name_len);
    const uint64_t cd_off = lh.size() + csize;
    if (cd_off >= 0xffffffffull) return GRAL_PF_EINVAL;
    const size_t cd_size = cd.size();
    put32(cd, 0x06054b50); put16(cd, 0); put16(cd, 0); put16(cd, 1); put16(cd, 1);
    put32(cd, (uint32_t)cd_size); put32(cd, (uint32_t)cd_off); put16(cd, 0);

    FILE *f = std::fopen(path, "wb");
    if (!f) return GRAL_PF_EINVAL;
    bool good = std::fwrite(lh.data(), 1, lh.size(), f) == lh.size();
    for (size_t c = 0; good && c < nch; ++c) good = std::fwrite(out[c].data(), 1, out[c].size(), f) == out[c].size();
    good = good && std::fwrite(cd.data(), 1, cd.size(), f) == cd.size();
    good = (std::fclose(f) == 0) && good;
    if (!good) {
        std::remove(path);                    // no corrupt file left behind (GFFWrite.cs:238-246)
        return GRAL_PF_EINVAL;
    }
    return GRAL_PF_OK;
}
