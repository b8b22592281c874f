/*
 * Minimal JSON DOM for the scene-file and glTF readers (RFC 8259). The
 * reference reads both formats with third-party parsers (cJSON for
 * src/app/scene/controller.c, cgltf for src/app/mesh/loader.c); this build
 * parses them with this reader instead.
 */
#ifndef VKRT_HOST_JSON_H
#define VKRT_HOST_JSON_H

#include <stddef.h>
#include <stdint.h>

typedef enum JsonType {
    JSON_NULL = 0,
    JSON_BOOL,
    JSON_NUMBER,
    JSON_STRING,
    JSON_ARRAY,
    JSON_OBJECT,
} JsonType;

typedef struct JsonValue {
    JsonType type;
    uint8_t boolean;
    double number;
    char* string;             /* JSON_STRING: decoded UTF-8, NUL-terminated */
    size_t stringLength;
    struct JsonValue* items;  /* JSON_ARRAY elements, JSON_OBJECT member values */
    char** keys;              /* JSON_OBJECT member names, parallel to items */
    uint32_t count;
} JsonValue;

/* Parses text[0..length). Returns 0 on success; on failure returns -1 and, when
 * error is given, writes a message with the byte offset of the problem. */
int vkrtJsonParse(const char* text, size_t length, JsonValue* out, char* error, size_t errorSize);
void vkrtJsonFree(JsonValue* value);

/* Member lookup (first match, case-sensitive); NULL when absent or not an object. */
const JsonValue* vkrtJsonGet(const JsonValue* object, const char* key);
const JsonValue* vkrtJsonAt(const JsonValue* array, uint32_t index);

int vkrtJsonIsNumber(const JsonValue* v);
/* Reads a non-negative integral number that fits in uint32. */
int vkrtJsonToU32(const JsonValue* v, uint32_t* out);
int vkrtJsonToFloat(const JsonValue* v, float* out);
/* Array of exactly count numbers. */
int vkrtJsonToFloats(const JsonValue* v, float* out, uint32_t count);

#endif
