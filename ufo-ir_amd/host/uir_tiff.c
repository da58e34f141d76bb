/*
 * uir_tiff.c -- baseline TIFF I/O for sinograms and slices (one float channel).
 *
 * What ufo-core's "read" / "write" nodes exchange with ufo-ir in a graph such
 * as json-sample.json:3-16: one 2-D frame per TIFF page.  Reading takes
 * uncompressed strips or tiles of 8-, 16- or 32-bit unsigned / signed integer
 * or 32- / 64-bit IEEE samples in either byte order, one sample per pixel
 * (extra samples are skipped), and converts them to FP32.  Writing produces
 * little-endian uncompressed 32-bit float pages (SampleFormat 3), one strip per
 * page, appended as further pages when asked.
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "uir_internal.h"

typedef struct {
    FILE *f;
    int be;          /* big-endian file */
    long size;
} tfile;

static uint16_t rd16(const tfile *t, const unsigned char *p)
{
    return t->be ? (uint16_t)(p[0] << 8 | p[1]) : (uint16_t)(p[1] << 8 | p[0]);
}

static uint32_t rd32(const tfile *t, const unsigned char *p)
{
    return t->be ? ((uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3])
                 : ((uint32_t)p[3] << 24 | (uint32_t)p[2] << 16 | (uint32_t)p[1] << 8 | p[0]);
}

static int rd_at(const tfile *t, long off, void *dst, size_t n)
{
    if (off < 0 || off + (long)n > t->size)
        return -1;
    if (fseek(t->f, off, SEEK_SET) != 0)
        return -1;
    return fread(dst, 1, n, t->f) == n ? 0 : -1;
}

static uir_status t_open(const char *path, tfile *t, uint32_t *first_ifd)
{
    memset(t, 0, sizeof *t);
    t->f = fopen(path, "rb");
    if (!t->f)
        return uir_set_error(UIR_ERR_NOT_FOUND, "cannot open %s", path);
    fseek(t->f, 0, SEEK_END);
    t->size = ftell(t->f);
    unsigned char h[8];
    if (rd_at(t, 0, h, 8) != 0) {
        fclose(t->f);
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "%s: not a TIFF file", path);
    }
    if (h[0] == 'I' && h[1] == 'I')
        t->be = 0;
    else if (h[0] == 'M' && h[1] == 'M')
        t->be = 1;
    else {
        fclose(t->f);
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "%s: not a TIFF file", path);
    }
    if (rd16(t, h + 2) != 42) {
        fclose(t->f);
        return uir_set_error(UIR_ERR_UNSUPPORTED, "%s: not a classic TIFF (BigTIFF is not supported)", path);
    }
    *first_ifd = rd32(t, h + 4);
    return UIR_OK;
}

typedef struct {
    uint32_t w, h, bits, fmt, spp, compression, planar;
    uint32_t rows_per_strip, tile_w, tile_h;
    uint32_t n_off, n_cnt;
    long off_pos, cnt_pos;   /* file offset of the offsets / counts arrays */
    uint16_t off_type, cnt_type;
    uint32_t next;
} tifd;

/* value of a SHORT / LONG entry with count 1 */
static uint32_t entry_val(const tfile *t, const unsigned char *e)
{
    const uint16_t type = rd16(t, e + 2);
    return type == 3 ? rd16(t, e + 8) : rd32(t, e + 8);
}

static uir_status read_ifd(const tfile *t, uint32_t off, tifd *d)
{
    memset(d, 0, sizeof *d);
    d->bits = 1;
    d->fmt = 1;
    d->spp = 1;
    d->compression = 1;
    d->planar = 1;
    d->rows_per_strip = 0xffffffffu;
    unsigned char nb[2];
    if (rd_at(t, off, nb, 2) != 0)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "TIFF: bad IFD offset");
    const uint16_t n = rd16(t, nb);
    unsigned char *buf = (unsigned char *)malloc((size_t)n * 12 + 4);
    if (!buf)
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
    if (rd_at(t, off + 2, buf, (size_t)n * 12 + 4) != 0) {
        free(buf);
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "TIFF: truncated IFD");
    }
    for (uint16_t i = 0; i < n; ++i) {
        const unsigned char *e = buf + 12 * i;
        const uint16_t tag = rd16(t, e), type = rd16(t, e + 2);
        const uint32_t count = rd32(t, e + 4);
        const size_t tsz = type == 3 ? 2 : 4;
        switch (tag) {
        case 256: d->w = entry_val(t, e); break;
        case 257: d->h = entry_val(t, e); break;
        case 258: /* BitsPerSample: the first sample's */
            if (type == 3 && count > 2) {
                unsigned char v[2];
                if (rd_at(t, rd32(t, e + 8), v, 2) == 0)
                    d->bits = rd16(t, v);
            } else {
                d->bits = type == 3 ? rd16(t, e + 8) : rd32(t, e + 8);
            }
            break;
        case 259: d->compression = entry_val(t, e); break;
        case 273:
        case 324: /* StripOffsets / TileOffsets */
            d->n_off = count;
            d->off_type = type;
            d->off_pos = count * tsz <= 4 ? (long)(off + 2 + 12 * i + 8) : (long)rd32(t, e + 8);
            break;
        case 277: d->spp = entry_val(t, e); break;
        case 278: d->rows_per_strip = entry_val(t, e); break;
        case 279:
        case 325: /* StripByteCounts / TileByteCounts */
            d->n_cnt = count;
            d->cnt_type = type;
            d->cnt_pos = count * tsz <= 4 ? (long)(off + 2 + 12 * i + 8) : (long)rd32(t, e + 8);
            break;
        case 284: d->planar = entry_val(t, e); break;
        case 322: d->tile_w = entry_val(t, e); break;
        case 323: d->tile_h = entry_val(t, e); break;
        case 339: d->fmt = entry_val(t, e); break;
        default: break;
        }
    }
    d->next = rd32(t, buf + (size_t)n * 12);
    free(buf);
    if (!d->w || !d->h || !d->n_off)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "TIFF: page without size or data");
    return UIR_OK;
}

static uint32_t arr_get(const tfile *t, long pos, uint16_t type, uint32_t i, int *err)
{
    unsigned char v[4];
    const size_t sz = type == 3 ? 2 : 4;
    if (rd_at(t, pos + (long)(i * sz), v, sz) != 0) {
        *err = 1;
        return 0;
    }
    return sz == 2 ? rd16(t, v) : rd32(t, v);
}

static float sample(const tfile *t, const unsigned char *p, uint32_t bits, uint32_t fmt)
{
    switch (bits) {
    case 8: return fmt == 2 ? (float)(int8_t)p[0] : (float)p[0];
    case 16: {
        const uint16_t v = rd16(t, p);
        return fmt == 2 ? (float)(int16_t)v : (float)v;
    }
    case 32: {
        const uint32_t v = rd32(t, p);
        if (fmt == 3) {
            float f;
            memcpy(&f, &v, 4);
            return f;
        }
        return fmt == 2 ? (float)(int32_t)v : (float)v;
    }
    default: { /* 64-bit float */
        uint64_t v = 0;
        for (int k = 0; k < 8; ++k)
            v |= (uint64_t)p[t->be ? k : 7 - k] << (8 * (7 - k));
        double d;
        memcpy(&d, &v, 8);
        return (float)d;
    }
    }
}

static uir_status seek_page(const tfile *t, uint32_t first, uint32_t page, tifd *d)
{
    uint32_t off = first;
    for (uint32_t p = 0;; ++p) {
        if (!off)
            return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "TIFF: page %u does not exist", page);
        uir_status s = read_ifd(t, off, d);
        if (s != UIR_OK)
            return s;
        if (p == page)
            return UIR_OK;
        off = d->next;
        if (p > 1000000)
            return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "TIFF: IFD loop");
    }
}

uir_status uir_tiff_info(const char *path, uint32_t *n_pages, uint32_t *width, uint32_t *height)
{
    if (!path)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL path");
    tfile t;
    uint32_t off;
    uir_status s = t_open(path, &t, &off);
    if (s != UIR_OK)
        return s;
    uint32_t np = 0;
    tifd d;
    while (off && s == UIR_OK) {
        s = read_ifd(&t, off, &d);
        if (s != UIR_OK)
            break;
        if (np == 0) {
            if (width) *width = d.w;
            if (height) *height = d.h;
        }
        ++np;
        off = d.next;
        if (np > 1000000)
            s = uir_set_error(UIR_ERR_INVALID_ARGUMENT, "TIFF: IFD loop");
    }
    fclose(t.f);
    if (s == UIR_OK && n_pages)
        *n_pages = np;
    return s;
}

uir_status uir_tiff_read(const char *path, uint32_t page, float *dst, size_t capacity, uint32_t *width,
                         uint32_t *height)
{
    if (!path)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL path");
    tfile t;
    uint32_t first;
    uir_status s = t_open(path, &t, &first);
    if (s != UIR_OK)
        return s;
    tifd d;
    if ((s = seek_page(&t, first, page, &d)) != UIR_OK)
        goto out;
    if (width) *width = d.w;
    if (height) *height = d.h;
    if (!dst)
        goto out;
    if ((size_t)d.w * d.h > capacity) {
        s = uir_set_error(UIR_ERR_INVALID_ARGUMENT, "TIFF page of %ux%u does not fit %zu floats", d.w, d.h, capacity);
        goto out;
    }
    if (d.compression != 1) {
        s = uir_set_error(UIR_ERR_UNSUPPORTED, "%s: compressed TIFF (scheme %u) is not supported", path, d.compression);
        goto out;
    }
    if (!((d.bits == 8 || d.bits == 16 || d.bits == 32) && (d.fmt == 1 || d.fmt == 2 || (d.fmt == 3 && d.bits == 32))) &&
        !(d.bits == 64 && d.fmt == 3)) {
        s = uir_set_error(UIR_ERR_UNSUPPORTED, "%s: %u-bit samples of format %u are not supported", path, d.bits, d.fmt);
        goto out;
    }
    if (d.spp > 1 && d.planar != 1) {
        s = uir_set_error(UIR_ERR_UNSUPPORTED, "%s: planar multi-sample TIFF is not supported", path);
        goto out;
    }
    if (d.w == 0 || d.h == 0) {
        s = uir_set_error(UIR_ERR_INVALID_ARGUMENT, "%s: empty page (%ux%u)", path, d.w, d.h);
        goto out;
    }
    if ((d.tile_w == 0) != (d.tile_h == 0)) {
        s = uir_set_error(UIR_ERR_INVALID_ARGUMENT, "%s: TileWidth %u with TileLength %u", path, d.tile_w, d.tile_h);
        goto out;
    }
    if (!d.tile_w && d.rows_per_strip == 0) {
        s = uir_set_error(UIR_ERR_INVALID_ARGUMENT, "%s: RowsPerStrip is 0", path);
        goto out;
    }
    {
        const size_t bps = d.bits / 8, px = bps * d.spp;
        const int tiled = d.tile_w && d.tile_h;
        const uint32_t bw = tiled ? d.tile_w : d.w;
        const uint32_t bh = tiled ? d.tile_h : (d.rows_per_strip > d.h ? d.h : d.rows_per_strip);
        const uint32_t across = tiled ? (d.w + bw - 1) / bw : 1;
        const uint32_t down = (d.h + bh - 1) / bh;
        if ((uint64_t)across * down > d.n_off) {
            s = uir_set_error(UIR_ERR_INVALID_ARGUMENT, "%s: %u data blocks, %u expected", path, d.n_off, across * down);
            goto out;
        }
        unsigned char *blk = (unsigned char *)malloc((size_t)bw * bh * px);
        if (!blk) {
            s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
            goto out;
        }
        int err = 0;
        for (uint32_t by = 0; by < down && s == UIR_OK; ++by)
            for (uint32_t bx = 0; bx < across && s == UIR_OK; ++bx) {
                const uint32_t k = by * across + bx;
                const uint32_t boff = arr_get(&t, d.off_pos, d.off_type, k, &err);
                const uint32_t rows =
                    tiled ? bh : ((uint64_t)(by + 1) * bh > d.h ? (uint32_t)(d.h - (uint64_t)by * bh) : bh);
                const size_t need = (size_t)bw * rows * px;
                if (err || rd_at(&t, boff, blk, need) != 0) {
                    s = uir_set_error(UIR_ERR_INVALID_ARGUMENT, "%s: truncated image data", path);
                    break;
                }
                for (uint32_t y = 0; y < rows; ++y) {
                    const uint64_t gy = (uint64_t)by * bh + y;
                    if (gy >= d.h)
                        break;
                    for (uint32_t x = 0; x < bw; ++x) {
                        const uint64_t gx = (uint64_t)bx * bw + x;
                        if (gx >= d.w)
                            break;
                        dst[(size_t)gy * d.w + gx] = sample(&t, blk + ((size_t)y * bw + x) * px, d.bits, d.fmt);
                    }
                }
            }
        free(blk);
    }
out:
    fclose(t.f);
    return s;
}

static void w16(unsigned char *p, uint16_t v) { p[0] = (unsigned char)v; p[1] = (unsigned char)(v >> 8); }
static void w32(unsigned char *p, uint32_t v)
{
    p[0] = (unsigned char)v;
    p[1] = (unsigned char)(v >> 8);
    p[2] = (unsigned char)(v >> 16);
    p[3] = (unsigned char)(v >> 24);
}

uir_status uir_tiff_write(const char *path, const float *data, uint32_t width, uint32_t height, int append)
{
    if (!path || !data || !width || !height)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "bad arguments");
    const uint64_t bytes = (uint64_t)width * height * 4;
    FILE *f = NULL;
    uint32_t link_pos = 4; /* where the offset of the new IFD goes */
    long end = 0;
    if (append) {
        f = fopen(path, "r+b");
        if (f) {
            tfile t;
            t.f = f;
            fseek(f, 0, SEEK_END);
            t.size = ftell(f);
            unsigned char h[8];
            if (rd_at(&t, 0, h, 8) != 0 || h[0] != 'I' || h[1] != 'I') {
                fclose(f);
                return uir_set_error(UIR_ERR_UNSUPPORTED, "%s: can only append to little-endian TIFF", path);
            }
            t.be = 0;
            uint32_t off = rd32(&t, h + 4);
            while (off) {
                unsigned char nb[2];
                if (rd_at(&t, off, nb, 2) != 0) {
                    fclose(f);
                    return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "%s: bad IFD chain", path);
                }
                link_pos = off + 2 + 12u * rd16(&t, nb);
                unsigned char nx[4];
                if (rd_at(&t, link_pos, nx, 4) != 0) {
                    fclose(f);
                    return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "%s: bad IFD chain", path);
                }
                off = rd32(&t, nx);
            }
            end = t.size;
        }
    }
    if (!f) {
        f = fopen(path, "wb");
        if (!f)
            return uir_set_error(UIR_ERR_IO, "cannot create %s", path);
        unsigned char h[8] = {'I', 'I', 42, 0, 0, 0, 0, 0};
        fwrite(h, 1, 8, f);
        end = 8;
    }
    if (end & 1) { /* word alignment */
        fseek(f, end, SEEK_SET);
        fputc(0, f);
        ++end;
    }
    if (bytes + (uint64_t)end + 256 >= (1ull << 32)) {
        fclose(f);
        return uir_set_error(UIR_ERR_UNSUPPORTED, "%s: classic TIFF is limited to 4 GiB", path);
    }
    /* layout: pixel data, then the IFD */
    const uint32_t data_off = (uint32_t)end;
    const uint32_t ifd_off = (uint32_t)(end + bytes);
    fseek(f, end, SEEK_SET);
    unsigned char *row = (unsigned char *)malloc((size_t)width * 4);
    if (!row) {
        fclose(f);
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
    }
    for (uint32_t y = 0; y < height; ++y) {
        for (uint32_t x = 0; x < width; ++x) {
            uint32_t v;
            memcpy(&v, &data[(size_t)y * width + x], 4);
            w32(row + 4 * x, v);
        }
        fwrite(row, 1, (size_t)width * 4, f);
    }
    free(row);
    static const uint16_t tags[] = {256, 257, 258, 259, 262, 273, 277, 278, 279, 284, 339};
    const uint32_t vals[] = {width, height, 32, 1, 1, data_off, 1, height, (uint32_t)bytes, 1, 3};
    const uint16_t types[] = {4, 4, 3, 3, 3, 4, 3, 4, 4, 3, 3};
    const int nt = (int)(sizeof tags / sizeof tags[0]);
    unsigned char ifd[2 + 12 * 11 + 4];
    w16(ifd, (uint16_t)nt);
    for (int i = 0; i < nt; ++i) {
        unsigned char *e = ifd + 2 + 12 * i;
        w16(e, tags[i]);
        w16(e + 2, types[i]);
        w32(e + 4, 1);
        memset(e + 8, 0, 4);
        if (types[i] == 3)
            w16(e + 8, (uint16_t)vals[i]);
        else
            w32(e + 8, vals[i]);
    }
    w32(ifd + 2 + 12 * nt, 0);
    fwrite(ifd, 1, sizeof ifd, f);
    unsigned char lk[4];
    w32(lk, ifd_off);
    fseek(f, link_pos, SEEK_SET);
    fwrite(lk, 1, 4, f);
    const int bad = ferror(f);
    fclose(f);
    if (bad)
        return uir_set_error(UIR_ERR_IO, "%s: write error", path);
    return UIR_OK;
}
