// Uncompressed TIFF output of the aligned-image path (writeAlignedImageTask, src/remap.zig:134-164).
//
// The reference sets its tags through libtiff (openTiffWriter, src/image_io.zig:100-143) and writes
// scanlines (writeTiff, :175-200).  fg_write_tiff produces the file that call sequence produces,
// byte for byte, without libtiff: a classic little-endian TIFF with
//   - the pixel data at offset 8, then the directory on the next even offset, then the directory's
//     out-of-line values in the order libtiff emits them: XResolution, YResolution, BitsPerSample,
//     StripByteCounts, StripOffsets, SampleFormat;
//   - RowsPerStrip = the row size in bytes: openTiffWriter passes the row size as the *request* of
//     TIFFDefaultStripSize, which returns any positive request unchanged, so images with fewer
//     rows than bytes per row are one strip;
//   - ImageWidth, ImageLength and RowsPerStrip as SHORT when they fit, else LONG; StripOffsets as
//     LONG; StripByteCounts as SHORT when there are several strips of at most 65535 bytes, else
//     LONG (libtiff's choices for uncompressed classic files).
#include <stdio.h>

#include <string>
#include <vector>

#include "fg_host.h"

namespace {

enum : uint16_t { T_SHORT = 3, T_LONG = 4, T_RATIONAL = 5 };

struct Tag {
    uint16_t tag, type;
    uint32_t count;
    std::vector<uint8_t> bytes;  // the value, little endian
};

void put16(std::vector<uint8_t>& b, uint32_t v) {
    b.push_back((uint8_t)v);
    b.push_back((uint8_t)(v >> 8));
}
void put32(std::vector<uint8_t>& b, uint32_t v) {
    for (int i = 0; i < 4; i++) b.push_back((uint8_t)(v >> (8 * i)));
}

Tag shorts(uint16_t tag, const std::vector<uint32_t>& v) {
    Tag t{tag, T_SHORT, (uint32_t)v.size(), {}};
    for (uint32_t x : v) put16(t.bytes, x);
    return t;
}
Tag longs(uint16_t tag, const std::vector<uint32_t>& v) {
    Tag t{tag, T_LONG, (uint32_t)v.size(), {}};
    for (uint32_t x : v) put32(t.bytes, x);
    return t;
}
Tag short_or_long(uint16_t tag, uint32_t v) { return v <= 0xFFFFu ? shorts(tag, {v}) : longs(tag, {v}); }
Tag rational(uint16_t tag, uint32_t num, uint32_t den) {
    Tag t{tag, T_RATIONAL, 1, {}};
    put32(t.bytes, num);
    put32(t.bytes, den);
    return t;
}

}  // namespace

extern "C" int fg_write_tiff(const char* path, const fg_image_desc* d, const void* px) {
    if (!path || !d || !px) return fg_set_error(FG_ERR_INVALID_ARG, "null argument");
    if (d->width == 0 || d->height == 0 || (d->color_channels != 1 && d->color_channels != 3) || d->extra_channels > 1 ||
        (d->sample_type != FG_SAMPLE_U8 && d->sample_type != FG_SAMPLE_U16))
        return fg_set_error(FG_ERR_INVALID_ARG, "unsupported image description");
    const uint32_t spp = d->color_channels + d->extra_channels;
    const uint32_t bps = d->sample_type == FG_SAMPLE_U16 ? 16 : 8;
    const uint64_t row_bytes = (uint64_t)d->width * spp * (bps / 8);
    const uint64_t data_bytes = row_bytes * d->height;
    // the directory and its values must stay addressable with 32-bit offsets
    if (row_bytes > 0xFFFFFFFFull || data_bytes > 0xFFFF0000ull)
        return fg_set_error(FG_ERR_UNSUPPORTED, "image too large for a classic TIFF");
    const uint32_t rows_per_strip = (uint32_t)row_bytes;
    const uint32_t nstrips = (uint32_t)((d->height + (uint64_t)rows_per_strip - 1) / rows_per_strip);
    const uint32_t data_off = 8;
    const uint32_t ifd_off = (uint32_t)((data_off + data_bytes + 1) & ~1ull);

    std::vector<uint32_t> strip_off(nstrips), strip_bytes(nstrips);
    for (uint32_t s = 0; s < nstrips; s++) {
        const uint64_t row0 = (uint64_t)s * rows_per_strip;
        const uint64_t rows = (d->height - row0 < rows_per_strip) ? d->height - row0 : rows_per_strip;
        strip_off[s] = (uint32_t)(data_off + row0 * row_bytes);
        strip_bytes[s] = (uint32_t)(rows * row_bytes);
    }
    const uint64_t full_strip = (uint64_t)(rows_per_strip < d->height ? rows_per_strip : d->height) * row_bytes;

    std::vector<Tag> tags;  // ascending tag order
    tags.push_back(short_or_long(256, d->width));                                   // ImageWidth
    tags.push_back(short_or_long(257, d->height));                                  // ImageLength
    tags.push_back(shorts(258, std::vector<uint32_t>(spp, bps)));                   // BitsPerSample
    tags.push_back(shorts(259, {1}));                                               // Compression: none
    tags.push_back(shorts(262, {d->color_channels == 3 ? 2u : 1u}));                // Photometric: RGB / MinIsBlack
    tags.push_back(longs(273, strip_off));                                          // StripOffsets
    tags.push_back(shorts(274, {1}));                                               // Orientation: top-left
    tags.push_back(shorts(277, {spp}));                                             // SamplesPerPixel
    tags.push_back(short_or_long(278, rows_per_strip));                             // RowsPerStrip
    tags.push_back(nstrips > 1 && full_strip <= 0xFFFFu ? shorts(279, strip_bytes)  // StripByteCounts
                                                        : longs(279, strip_bytes));
    tags.push_back(rational(282, 150, 1));                                          // XResolution
    tags.push_back(rational(283, 150, 1));                                          // YResolution
    tags.push_back(shorts(284, {1}));                                               // PlanarConfiguration: contig
    tags.push_back(shorts(296, {2}));                                               // ResolutionUnit: inch
    if (d->extra_channels) tags.push_back(shorts(338, {2}));                        // ExtraSamples: unassoc. alpha
    tags.push_back(shorts(339, std::vector<uint32_t>(spp, 1)));                     // SampleFormat: uint
    tags.push_back(longs(33300, {d->width}));                                       // PixarImageFullWidth
    tags.push_back(longs(33301, {d->height}));                                      // PixarImageFullLength

    // out-of-line values follow the directory, in libtiff's emission order
    const uint32_t ool_base = ifd_off + 2 + 12 * (uint32_t)tags.size() + 4;
    if ((uint64_t)ool_base + 16 + 8ull * spp + 8ull * nstrips > 0xFFFFFFFFull)
        return fg_set_error(FG_ERR_UNSUPPORTED, "image too large for a classic TIFF");
    std::vector<uint8_t> ool;
    std::vector<uint32_t> value(tags.size());
    for (uint16_t want : {282, 283, 258, 279, 273, 339}) {
        for (size_t i = 0; i < tags.size(); i++) {
            if (tags[i].tag != want || tags[i].bytes.size() <= 4) continue;
            value[i] = ool_base + (uint32_t)ool.size();
            ool.insert(ool.end(), tags[i].bytes.begin(), tags[i].bytes.end());
        }
    }
    std::vector<uint8_t> ifd;
    put16(ifd, (uint32_t)tags.size());
    for (size_t i = 0; i < tags.size(); i++) {
        const Tag& t = tags[i];
        put16(ifd, t.tag);
        put16(ifd, t.type);
        put32(ifd, t.count);
        if (t.bytes.size() <= 4) {  // in the entry, left-justified
            std::vector<uint8_t> v = t.bytes;
            v.resize(4, 0);
            ifd.insert(ifd.end(), v.begin(), v.end());
        } else {
            put32(ifd, value[i]);
        }
    }
    put32(ifd, 0);  // no next directory

    std::vector<uint8_t> head{'I', 'I'};
    put16(head, 42);
    put32(head, ifd_off);
    FILE* f = fopen(path, "wb");
    if (!f) return fg_set_error(FG_ERR_INVALID_ARG, (std::string("cannot open ") + path).c_str());
    const uint8_t pad = 0;
    bool ok = fwrite(head.data(), 1, head.size(), f) == head.size();
    // samples in host order, which is little endian, as the header says
    ok = ok && fwrite(px, 1, (size_t)data_bytes, f) == (size_t)data_bytes;
    if (ok && (data_bytes & 1)) ok = fwrite(&pad, 1, 1, f) == 1;
    ok = ok && fwrite(ifd.data(), 1, ifd.size(), f) == ifd.size();
    ok = ok && fwrite(ool.data(), 1, ool.size(), f) == ool.size();
    ok = (fclose(f) == 0) && ok;
    if (!ok) return fg_set_error(FG_ERR_DEVICE, (std::string("write failed: ") + path).c_str());
    return FG_OK;
}

// writeAlignedImageTask (src/remap.zig:134-164) for one frame already in memory: remapRigidImage
// into the ROI (colour channels plus the support alpha), then the TIFF "{prefix}_{index:04}.tif".
extern "C" int fg_write_aligned_tiff(fg_context* ctx, const fg_image_desc* src, const void* src_px, const fg_pose* pose,
                                     const fg_rect* roi, const char* prefix, uint32_t output_index) {
    if (!ctx || !src || !src_px || !pose || !prefix) return fg_set_error(FG_ERR_INVALID_ARG, "null argument");
    const fg_rect r = roi ? *roi : fg_rect{0, 0, (int32_t)src->width, (int32_t)src->height};
    if (r.right <= r.left || r.bottom <= r.top) return fg_set_error(FG_ERR_INVALID_ARG, "empty roi");
    fg_image_desc od = *src;
    od.width = (uint32_t)(r.right - r.left);
    od.height = (uint32_t)(r.bottom - r.top);
    od.extra_channels = 1;
    const size_t bytes =
        (size_t)od.width * od.height * (od.color_channels + 1) * (od.sample_type == FG_SAMPLE_U16 ? 2 : 1);
    std::vector<uint8_t> out;
    try {
        out.resize(bytes);
    } catch (...) {
        return fg_set_error(FG_ERR_OOM_HOST, "host allocation failed");
    }
    int rc = fg_remap(ctx, src, src_px, pose, &r, out.data());
    if (rc) return rc;
    std::vector<char> path(std::string(prefix).size() + 32);
    snprintf(path.data(), path.size(), "%s_%04u.tif", prefix, output_index);
    return fg_write_tiff(path.data(), &od, out.data());
}
