/*
 * Recursive-descent JSON reader. Strings are decoded (escapes, \u surrogate
 * pairs to UTF-8); numbers follow the JSON grammar and are converted with
 * strtod on a bounded copy; nesting is limited so hostile input cannot exhaust
 * the stack.
 */
#include "json.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { kJsonMaxDepth = 256 };

typedef struct Parser {
    const char* text;
    size_t length;
    size_t pos;
    char* error;
    size_t errorSize;
    int failed;
} Parser;

static int fail(Parser* p, const char* fmt, ...) {
    if (!p->failed && p->error && p->errorSize) {
        char msg[160];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(msg, sizeof(msg), fmt, ap);
        va_end(ap);
        snprintf(p->error, p->errorSize, "JSON error at byte %zu: %s", p->pos, msg);
    }
    p->failed = 1;
    return -1;
}

static void skipSpace(Parser* p) {
    while (p->pos < p->length) {
        char c = p->text[p->pos];
        if (c != ' ' && c != '\t' && c != '\n' && c != '\r') break;
        p->pos++;
    }
}

static int parseValue(Parser* p, JsonValue* out, int depth);

static int hexDigit(char c) {
    if (c >= '0' && c <= '9') return c - '0';
    if (c >= 'a' && c <= 'f') return c - 'a' + 10;
    if (c >= 'A' && c <= 'F') return c - 'A' + 10;
    return -1;
}

static int readHex4(Parser* p, uint32_t* out) {
    if (p->pos + 4 > p->length) return fail(p, "truncated \\u escape");
    uint32_t v = 0;
    for (int i = 0; i < 4; i++) {
        int d = hexDigit(p->text[p->pos + (size_t)i]);
        if (d < 0) return fail(p, "bad \\u escape");
        v = (v << 4) | (uint32_t)d;
    }
    p->pos += 4;
    *out = v;
    return 0;
}

static size_t putUtf8(char* dst, uint32_t cp) {
    if (cp < 0x80u) {
        dst[0] = (char)cp;
        return 1;
    }
    if (cp < 0x800u) {
        dst[0] = (char)(0xC0u | (cp >> 6));
        dst[1] = (char)(0x80u | (cp & 0x3Fu));
        return 2;
    }
    if (cp < 0x10000u) {
        dst[0] = (char)(0xE0u | (cp >> 12));
        dst[1] = (char)(0x80u | ((cp >> 6) & 0x3Fu));
        dst[2] = (char)(0x80u | (cp & 0x3Fu));
        return 3;
    }
    dst[0] = (char)(0xF0u | (cp >> 18));
    dst[1] = (char)(0x80u | ((cp >> 12) & 0x3Fu));
    dst[2] = (char)(0x80u | ((cp >> 6) & 0x3Fu));
    dst[3] = (char)(0x80u | (cp & 0x3Fu));
    return 4;
}

/* p->pos is just past the opening quote. The decoded text is never longer
 * than the escaped source, so one allocation of the source span suffices. */
static int parseString(Parser* p, char** outString, size_t* outLength) {
    size_t start = p->pos;
    size_t end = start;
    while (end < p->length && p->text[end] != '"') {
        if (p->text[end] == '\\') end++;
        end++;
    }
    if (end >= p->length) return fail(p, "unterminated string");
    char* s = (char*)malloc(end - start + 1);
    if (!s) return fail(p, "out of memory");
    size_t n = 0;
    while (p->pos < end) {
        unsigned char c = (unsigned char)p->text[p->pos++];
        if (c < 0x20u) {
            free(s);
            return fail(p, "control character in string");
        }
        if (c != '\\') {
            s[n++] = (char)c;
            continue;
        }
        char e = p->text[p->pos++];
        switch (e) {
            case '"': s[n++] = '"'; break;
            case '\\': s[n++] = '\\'; break;
            case '/': s[n++] = '/'; break;
            case 'b': s[n++] = '\b'; break;
            case 'f': s[n++] = '\f'; break;
            case 'n': s[n++] = '\n'; break;
            case 'r': s[n++] = '\r'; break;
            case 't': s[n++] = '\t'; break;
            case 'u': {
                uint32_t cp = 0;
                if (readHex4(p, &cp) != 0) {
                    free(s);
                    return -1;
                }
                if (cp >= 0xD800u && cp <= 0xDBFFu) {
                    uint32_t lo = 0;
                    if (p->pos + 2 > end || p->text[p->pos] != '\\' || p->text[p->pos + 1] != 'u') {
                        free(s);
                        return fail(p, "unpaired surrogate");
                    }
                    p->pos += 2;
                    if (readHex4(p, &lo) != 0 || lo < 0xDC00u || lo > 0xDFFFu) {
                        free(s);
                        return fail(p, "bad surrogate pair");
                    }
                    cp = 0x10000u + ((cp - 0xD800u) << 10) + (lo - 0xDC00u);
                } else if (cp >= 0xDC00u && cp <= 0xDFFFu) {
                    free(s);
                    return fail(p, "unpaired surrogate");
                }
                n += putUtf8(s + n, cp);
                break;
            }
            default:
                free(s);
                return fail(p, "bad escape");
        }
    }
    p->pos = end + 1;
    s[n] = '\0';
    *outString = s;
    *outLength = n;
    return 0;
}

static int parseNumber(Parser* p, JsonValue* out) {
    size_t start = p->pos;
    size_t i = p->pos;
    const char* t = p->text;
    if (i < p->length && t[i] == '-') i++;
    if (i >= p->length) return fail(p, "bad number");
    if (t[i] == '0') {
        i++;
    } else if (t[i] >= '1' && t[i] <= '9') {
        while (i < p->length && t[i] >= '0' && t[i] <= '9') i++;
    } else {
        return fail(p, "bad number");
    }
    if (i < p->length && t[i] == '.') {
        i++;
        if (i >= p->length || t[i] < '0' || t[i] > '9') return fail(p, "bad fraction");
        while (i < p->length && t[i] >= '0' && t[i] <= '9') i++;
    }
    if (i < p->length && (t[i] == 'e' || t[i] == 'E')) {
        i++;
        if (i < p->length && (t[i] == '+' || t[i] == '-')) i++;
        if (i >= p->length || t[i] < '0' || t[i] > '9') return fail(p, "bad exponent");
        while (i < p->length && t[i] >= '0' && t[i] <= '9') i++;
    }
    char buf[64];
    size_t len = i - start;
    char* heap = NULL;
    char* src = buf;
    if (len >= sizeof(buf)) {
        heap = (char*)malloc(len + 1);
        if (!heap) return fail(p, "out of memory");
        src = heap;
    }
    memcpy(src, t + start, len);
    src[len] = '\0';
    out->type = JSON_NUMBER;
    out->number = strtod(src, NULL);
    free(heap);
    p->pos = i;
    return 0;
}

static int matchWord(Parser* p, const char* word) {
    size_t n = strlen(word);
    if (p->pos + n > p->length || memcmp(p->text + p->pos, word, n) != 0) return 0;
    p->pos += n;
    return 1;
}

static int growItems(JsonValue* v, uint32_t* capacity, int withKeys) {
    if (v->count < *capacity) return 0;
    uint32_t cap = *capacity ? *capacity * 2u : 4u;
    JsonValue* items = (JsonValue*)realloc(v->items, sizeof(JsonValue) * cap);
    if (!items) return -1;
    v->items = items;
    if (withKeys) {
        char** keys = (char**)realloc(v->keys, sizeof(char*) * cap);
        if (!keys) return -1;
        v->keys = keys;
    }
    *capacity = cap;
    return 0;
}

static int parseArray(Parser* p, JsonValue* out, int depth) {
    out->type = JSON_ARRAY;
    uint32_t capacity = 0;
    skipSpace(p);
    if (p->pos < p->length && p->text[p->pos] == ']') {
        p->pos++;
        return 0;
    }
    for (;;) {
        if (growItems(out, &capacity, 0) != 0) return fail(p, "out of memory");
        JsonValue* item = &out->items[out->count];
        memset(item, 0, sizeof(*item));
        out->count++;
        if (parseValue(p, item, depth + 1) != 0) return -1;
        skipSpace(p);
        if (p->pos >= p->length) return fail(p, "unterminated array");
        char c = p->text[p->pos++];
        if (c == ']') return 0;
        if (c != ',') return fail(p, "expected ',' or ']'");
    }
}

static int parseObject(Parser* p, JsonValue* out, int depth) {
    out->type = JSON_OBJECT;
    uint32_t capacity = 0;
    skipSpace(p);
    if (p->pos < p->length && p->text[p->pos] == '}') {
        p->pos++;
        return 0;
    }
    for (;;) {
        skipSpace(p);
        if (p->pos >= p->length || p->text[p->pos] != '"') return fail(p, "expected member name");
        p->pos++;
        if (growItems(out, &capacity, 1) != 0) return fail(p, "out of memory");
        JsonValue* item = &out->items[out->count];
        memset(item, 0, sizeof(*item));
        out->keys[out->count] = NULL;
        out->count++;
        size_t keyLength = 0;
        if (parseString(p, &out->keys[out->count - 1], &keyLength) != 0) return -1;
        skipSpace(p);
        if (p->pos >= p->length || p->text[p->pos] != ':') return fail(p, "expected ':'");
        p->pos++;
        if (parseValue(p, item, depth + 1) != 0) return -1;
        skipSpace(p);
        if (p->pos >= p->length) return fail(p, "unterminated object");
        char c = p->text[p->pos++];
        if (c == '}') return 0;
        if (c != ',') return fail(p, "expected ',' or '}'");
    }
}

static int parseValue(Parser* p, JsonValue* out, int depth) {
    if (depth > kJsonMaxDepth) return fail(p, "nesting deeper than %d", kJsonMaxDepth);
    skipSpace(p);
    if (p->pos >= p->length) return fail(p, "unexpected end of input");
    char c = p->text[p->pos];
    switch (c) {
        case '{': p->pos++; return parseObject(p, out, depth);
        case '[': p->pos++; return parseArray(p, out, depth);
        case '"':
            p->pos++;
            out->type = JSON_STRING;
            return parseString(p, &out->string, &out->stringLength);
        case 't':
            if (!matchWord(p, "true")) return fail(p, "bad literal");
            out->type = JSON_BOOL;
            out->boolean = 1;
            return 0;
        case 'f':
            if (!matchWord(p, "false")) return fail(p, "bad literal");
            out->type = JSON_BOOL;
            out->boolean = 0;
            return 0;
        case 'n':
            if (!matchWord(p, "null")) return fail(p, "bad literal");
            out->type = JSON_NULL;
            return 0;
        default:
            if (c == '-' || (c >= '0' && c <= '9')) return parseNumber(p, out);
            return fail(p, "unexpected character 0x%02x", (unsigned)(unsigned char)c);
    }
}

int vkrtJsonParse(const char* text, size_t length, JsonValue* out, char* error, size_t errorSize) {
    if (!out) return -1;
    memset(out, 0, sizeof(*out));
    if (error && errorSize) error[0] = '\0';
    if (!text) return -1;
    Parser p = {text, length, 0, error, errorSize, 0};
    /* a UTF-8 byte-order mark is tolerated, as glTF's JSON chunk allows */
    if (length >= 3 && (unsigned char)text[0] == 0xEF && (unsigned char)text[1] == 0xBB &&
        (unsigned char)text[2] == 0xBF)
        p.pos = 3;
    if (parseValue(&p, out, 0) != 0) {
        vkrtJsonFree(out);
        return -1;
    }
    skipSpace(&p);
    /* GLB pads its JSON chunk with spaces; NULs after the document are tolerated as padding too */
    while (p.pos < p.length && p.text[p.pos] == '\0') p.pos++;
    if (p.pos != p.length) {
        fail(&p, "trailing characters");
        vkrtJsonFree(out);
        return -1;
    }
    return 0;
}

void vkrtJsonFree(JsonValue* v) {
    if (!v) return;
    for (uint32_t i = 0; i < v->count; i++) {
        if (v->items) vkrtJsonFree(&v->items[i]);
        if (v->keys) free(v->keys[i]);
    }
    free(v->items);
    free(v->keys);
    free(v->string);
    memset(v, 0, sizeof(*v));
}

const JsonValue* vkrtJsonGet(const JsonValue* object, const char* key) {
    if (!object || object->type != JSON_OBJECT || !key) return NULL;
    for (uint32_t i = 0; i < object->count; i++) {
        if (object->keys[i] && strcmp(object->keys[i], key) == 0) return &object->items[i];
    }
    return NULL;
}

const JsonValue* vkrtJsonAt(const JsonValue* array, uint32_t index) {
    if (!array || array->type != JSON_ARRAY || index >= array->count) return NULL;
    return &array->items[index];
}

int vkrtJsonIsNumber(const JsonValue* v) { return v && v->type == JSON_NUMBER; }

int vkrtJsonToU32(const JsonValue* v, uint32_t* out) {
    if (out) *out = 0;
    if (!vkrtJsonIsNumber(v) || !out) return 0;
    double d = v->number;
    if (!(d >= 0.0) || d > 4294967295.0) return 0;
    uint32_t u = (uint32_t)d;
    if ((double)u != d) return 0;
    *out = u;
    return 1;
}

int vkrtJsonToFloat(const JsonValue* v, float* out) {
    if (out) *out = 0.0f;
    if (!vkrtJsonIsNumber(v) || !out) return 0;
    *out = (float)v->number;
    return 1;
}

int vkrtJsonToFloats(const JsonValue* v, float* out, uint32_t count) {
    if (!v || v->type != JSON_ARRAY || v->count != count || !out) return 0;
    for (uint32_t i = 0; i < count; i++) {
        if (!vkrtJsonToFloat(&v->items[i], &out[i])) return 0;
    }
    return 1;
}
