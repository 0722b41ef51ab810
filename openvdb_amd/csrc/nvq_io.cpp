// .nvdb files (uncompressed, codec NONE) and raw grid-buffer files: read one grid into a host
// buffer for nvq_grid_upload, and write a single-grid file.  Host-only C++.
//
// File layout (nanovdb/nanovdb/io/IO.h:156-198, NanoVDB.h:5885-5927): a file is a sequence of
// segments; a segment is a 16-byte FileHeader {magic, version, gridCount, codec}, then per grid a
// 176-byte FileMetaData followed by the grid name (nameSize bytes including '\0'), then the grids'
// data, fileSize bytes each.  A raw file (NanoVDB.h:135-141, GridHandle::read) is grid buffers
// back to back, each mGridSize bytes.
#include "../../include/nvq.h"
#include "../../include/nvq_build.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace {

constexpr uint64_t kMagicNumb = 0x304244566f6e614eULL; // "NanoVDB0": grids and files before the split
constexpr uint64_t kMagicGrid = 0x314244566f6e614eULL; // "NanoVDB1": grid buffers
constexpr uint64_t kMagicFile = 0x324244566f6e614eULL; // "NanoVDB2": files
constexpr uint32_t kVersion = (32u << 21) | (9u << 10) | 2u; // Version() of the format this library writes
constexpr size_t kGridData = 672, kTreeData = 64, kFileHeader = 16, kFileMeta = 176;

template<typename T> T rd(const uint8_t* p)
{
    T v;
    std::memcpy(&v, p, sizeof(T));
    return v;
}
template<typename T> void wr(uint8_t* p, T v) { std::memcpy(p, &v, sizeof(T)); }

bool compatible(uint32_t version) { return (version >> 21) == 32u; } // Version::isCompatible: same major

// Reflected CRC-32 with polynomial 0xEDB88320, initial and final complement (the IEEE 802.3 /
// zlib CRC; tools/GridChecksum.h:117-126 computes the same function bit by bit).
uint32_t crc32(const uint8_t* p, size_t n)
{
    static const std::vector<uint32_t> table = [] {
        std::vector<uint32_t> t(256);
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c >> 1) ^ ((c & 1u) ? 0xEDB88320u : 0u);
            t[i] = c;
        }
        return t;
    }();
    uint32_t c = ~0u;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xffu] ^ (c >> 8);
    return ~c;
}

// io::stringHash (io/IO.h:718-729)
uint64_t nameHash(const char* s)
{
    uint64_t h = 0;
    for (auto* c = reinterpret_cast<const unsigned char*>(s); *c; ++c) {
        const uint64_t overflow = h >> 56;
        h *= 67;
        h += *c + overflow;
    }
    return h;
}

// tools::updateGridCount (tools/GridChecksum.h:412-420): grid index/count in the header, and the
// header CRC (crc32Head: bytes [16, GridData + TreeData), :224-230) when a checksum is present.
void setGridCount(uint8_t* g, uint32_t index, uint32_t count)
{
    if (rd<uint32_t>(g + 24) == index && rd<uint32_t>(g + 28) == count) return;
    wr<uint32_t>(g + 24, index);
    wr<uint32_t>(g + 28, count);
    if (rd<uint64_t>(g + 8) != ~uint64_t(0)) wr<uint32_t>(g + 8, crc32(g + 16, kGridData + kTreeData - 16));
}

struct File {
    FILE* f = nullptr;
    ~File()
    {
        if (f) std::fclose(f);
    }
    uint64_t size = 0; // file length in bytes
    bool open(const char* path)
    {
        f = std::fopen(path, "rb");
        if (!f || std::fseek(f, 0, SEEK_END) != 0) return false;
        const long end = std::ftell(f);
        if (end < 0) return false;
        size = uint64_t(end);
        return true;
    }
    // [pos, pos + n) lies inside the file (sizes read from the file are checked before use)
    bool holds(uint64_t pos, uint64_t n) const { return pos <= size && n <= size - pos; }
    bool read(uint64_t pos, void* dst, uint64_t n)
    {
        return holds(pos, n) && std::fseek(f, long(pos), SEEK_SET) == 0 && std::fread(dst, 1, n, f) == n;
    }
};

int take(File& F, uint64_t pos, uint64_t size, void** out, uint64_t* bytes)
{
    if (size < kGridData + kTreeData || !F.holds(pos, size)) return NVQ_ERR_TRUNCATED;
    uint8_t* buf = static_cast<uint8_t*>(std::aligned_alloc(32, (size + 31) & ~uint64_t(31)));
    if (!buf) return NVQ_ERR_INVALID_ARG;
    if (!F.read(pos, buf, size)) {
        std::free(buf);
        return NVQ_ERR_TRUNCATED;
    }
    setGridCount(buf, 0u, 1u);
    *out = buf;
    *bytes = size;
    return NVQ_OK;
}

bool nameMatches(const uint8_t* gridData, const char* name)
{
    char field[257];
    std::memcpy(field, gridData + 40, 256); // GridData::mGridName
    field[256] = '\0';
    return std::strcmp(field, name) == 0;
}

} // namespace

extern "C" {

uint32_t nvqb_crc32(const void* data, uint64_t size) { return crc32(static_cast<const uint8_t*>(data), size_t(size)); }

int nvqb_read_nvdb(const char* path, int32_t index, const char* name, void** out, uint64_t* bytes)
{
    if (!path || !out || !bytes || (!name && index < 0)) return NVQ_ERR_INVALID_ARG;
    *out = nullptr;
    *bytes = 0;
    File F;
    if (!F.open(path)) return NVQ_ERR_INVALID_ARG;
    uint8_t head[24];
    if (!F.read(0, head, sizeof(head))) return NVQ_ERR_TRUNCATED;
    const uint64_t magic = rd<uint64_t>(head);
    // raw grid buffers: "NanoVDB1", or "NanoVDB0" followed by a grid (GridData::mVersion at 16)
    const bool raw = magic == kMagicGrid || (magic == kMagicNumb && compatible(rd<uint32_t>(head + 16)));
    if (raw) {
        uint64_t pos = 0;
        for (int32_t k = 0;; ++k) {
            uint8_t g[kGridData];
            if (!F.read(pos, g, sizeof(g))) return NVQ_ERR_INVALID_ARG; // end of file: no such grid
            const uint64_t m = rd<uint64_t>(g);
            if ((m != kMagicGrid && m != kMagicNumb) || !compatible(rd<uint32_t>(g + 16))) return NVQ_ERR_BAD_MAGIC;
            const uint64_t size = rd<uint64_t>(g + 32); // GridData::mGridSize
            // a grid smaller than its own headers, or running past the end of the file, is corrupt
            // (and a zero size would never advance)
            if (size < kGridData + kTreeData || !F.holds(pos, size)) return NVQ_ERR_TRUNCATED;
            if (name ? nameMatches(g, name) : k == index) return take(F, pos, size, out, bytes);
            if (k == INT32_MAX) return NVQ_ERR_INVALID_ARG;
            pos += size;
        }
    }
    if (magic != kMagicFile && magic != kMagicNumb) return NVQ_ERR_BAD_MAGIC;
    const uint64_t key = name ? nameHash(name) : 0;
    uint64_t pos = 0;
    int32_t counter = -1;
    for (;;) {
        uint8_t fh[kFileHeader];
        if (!F.read(pos, fh, sizeof(fh))) return NVQ_ERR_INVALID_ARG; // end of file: no such grid
        const uint64_t m = rd<uint64_t>(fh);
        if (m != kMagicFile && m != kMagicNumb) return NVQ_ERR_BAD_MAGIC;
        if (!compatible(rd<uint32_t>(fh + 8))) return NVQ_ERR_BAD_VERSION;
        const uint16_t count = rd<uint16_t>(fh + 12), codec = rd<uint16_t>(fh + 14);
        pos += kFileHeader;
        struct Meta {
            uint64_t gridSize, fileSize, nameKey;
            std::string name;
        };
        std::vector<Meta> metas(count);
        for (auto& me : metas) {
            uint8_t fm[kFileMeta];
            if (!F.read(pos, fm, sizeof(fm))) return NVQ_ERR_TRUNCATED;
            me.gridSize = rd<uint64_t>(fm);
            me.fileSize = rd<uint64_t>(fm + 8);
            me.nameKey = rd<uint64_t>(fm + 16);
            const uint32_t nameSize = rd<uint32_t>(fm + 136);
            pos += kFileMeta;
            // sizes come from the file: the name must lie inside it (checked in 64 bits, so a
            // nameSize of 0xFFFFFFFF cannot wrap the buffer size), and so must the grid
            if (!F.holds(pos, nameSize) || !F.holds(pos + nameSize, 0)) return NVQ_ERR_TRUNCATED;
            std::vector<char> nm(size_t(nameSize) + 1, '\0');
            if (nameSize && !F.read(pos, nm.data(), nameSize)) return NVQ_ERR_TRUNCATED;
            me.name.assign(nm.data());
            pos += nameSize;
        }
        for (const auto& me : metas) {
            const bool hit = name ? ((me.nameKey == 0 || me.nameKey == key) && me.name == name) : ++counter == index;
            if (hit) {
                if (codec != 0) return NVQ_ERR_UNSUPPORTED_TYPE; // ZIP / BLOSC segments are not decoded here
                return take(F, pos, me.gridSize, out, bytes);
            }
            if (!F.holds(pos, me.fileSize)) return NVQ_ERR_TRUNCATED;
            pos += me.fileSize;
        }
    }
}

int nvqb_write_nvdb(const char* path, const void* grid_buf, uint64_t bytes)
{
    if (!path || !grid_buf || bytes < kGridData + kTreeData) return NVQ_ERR_INVALID_ARG;
    const uint8_t* g = static_cast<const uint8_t*>(grid_buf);
    const uint64_t m = rd<uint64_t>(g);
    if (m != kMagicGrid && m != kMagicNumb) return NVQ_ERR_BAD_MAGIC;
    if (!compatible(rd<uint32_t>(g + 16))) return NVQ_ERR_BAD_VERSION;
    const uint64_t gridSize = rd<uint64_t>(g + 32);
    if (gridSize > bytes) return NVQ_ERR_TRUNCATED;
    char gridName[257];
    std::memcpy(gridName, g + 40, 256);
    gridName[256] = '\0';
    const uint32_t nameSize = uint32_t(std::strlen(gridName) + 1);
    const uint8_t* tree = g + kGridData;
    const int64_t rootOff = rd<int64_t>(tree + 24);
    // io::writeGrid(fileName, handle, Codec::NONE) for a one-grid handle (io/IO.h:444-457, 321-351):
    uint8_t fh[kFileHeader] = {};
    wr<uint64_t>(fh, kMagicFile);
    wr<uint32_t>(fh + 8, kVersion);
    wr<uint16_t>(fh + 12, 1);
    wr<uint16_t>(fh + 14, 0); // Codec::NONE
    uint8_t fm[kFileMeta] = {};
    wr<uint64_t>(fm, gridSize);                          // gridSize
    wr<uint64_t>(fm + 8, gridSize);                      // fileSize: uncompressed
    wr<uint64_t>(fm + 16, nameHash(gridName));           // nameKey
    wr<uint64_t>(fm + 24, rd<uint64_t>(tree + 56));      // voxelCount (TreeData::mVoxelCount)
    std::memcpy(fm + 32, g + 636, 4);                    // gridType
    std::memcpy(fm + 36, g + 632, 4);                    // gridClass
    std::memcpy(fm + 40, g + 560, 48);                   // worldBBox
    std::memcpy(fm + 88, tree + rootOff, 24);            // indexBBox (RootData::mBBox)
    std::memcpy(fm + 112, g + 608, 24);                  // voxelSize
    wr<uint32_t>(fm + 136, nameSize);                    // nameSize
    std::memcpy(fm + 140, tree + 32, 12);                // nodeCount[0..2]
    wr<uint32_t>(fm + 152, 1u);                          // nodeCount[3] (the root)
    std::memcpy(fm + 156, tree + 44, 12);                // tileCount[0..2]
    wr<uint16_t>(fm + 168, 0);                           // codec NONE
    wr<uint16_t>(fm + 170, uint16_t(rd<uint32_t>(g + 648))); // blindDataCount
    wr<uint32_t>(fm + 172, kVersion);                    // version
    FILE* f = std::fopen(path, "wb");
    if (!f) return NVQ_ERR_INVALID_ARG;
    bool ok = std::fwrite(fh, 1, sizeof(fh), f) == sizeof(fh) && std::fwrite(fm, 1, sizeof(fm), f) == sizeof(fm) &&
              std::fwrite(gridName, 1, nameSize, f) == nameSize;
    if (ok && (rd<uint32_t>(g + 24) != 0 || rd<uint32_t>(g + 28) != 1)) {
        // a grid of a multi-grid buffer is written as grid 0 of 1 (writeUncompressedGrid, NanoVDB.h:5965-5972)
        std::unique_ptr<uint8_t[]> head(new uint8_t[kGridData]);
        std::memcpy(head.get(), g, kGridData);
        wr<uint32_t>(head.get() + 24, 0u);
        wr<uint32_t>(head.get() + 28, 1u);
        ok = std::fwrite(head.get(), 1, kGridData, f) == kGridData &&
             std::fwrite(g + kGridData, 1, gridSize - kGridData, f) == gridSize - kGridData;
    } else if (ok) {
        ok = std::fwrite(g, 1, gridSize, f) == gridSize;
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? NVQ_OK : NVQ_ERR_INVALID_ARG;
}

} // extern "C"
