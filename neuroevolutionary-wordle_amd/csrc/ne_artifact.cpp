// ne_artifact.cpp -- the winner artifact: a run's best genome as raw slot bytes plus JSON metadata with the
// action-space words (host only).  The JSON carries the reference's keys and format_version 2
// (src/model_artifact/winner_artifact.cpp:149-182) and one more key, "genome_layout": "FP32", because the
// bytes are this build's fp32 slot (the reference writes fp16 parameters); the reader refuses artifacts
// without it.
#include <cerrno>
#include <charconv>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <string>
#include <string_view>
#include <system_error>
#include <type_traits>
#include <vector>

#include "../../include/ne_runtime.h"
#include "ne_device.hpp"

namespace {

constexpr int kFormatVersion = 2;            // kSupportedWinnerArtifactFormatVersion (winner_artifact_reader.cpp:14)
constexpr uint64_t kCatalogCapacity = 4739;  // kTrainingWordCatalogCapacity (training_data.hpp:22)
constexpr const char* kGenomeLayout = "FP32";

std::string Escape(std::string_view in) {
    std::string out;
    for (const char c : in) {
        switch (c) {
            case '\\': out += "\\\\"; break;
            case '"': out += "\\\""; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            default:
                if (static_cast<unsigned char>(c) < 0x20) {  // other control bytes: \u00XX (RFC 8259 7)
                    static const char kHex[] = "0123456789abcdef";
                    out += "\\u00";
                    out += kHex[static_cast<unsigned char>(c) >> 4];
                    out += kHex[static_cast<unsigned char>(c) & 15];
                } else {
                    out += c;
                }
        }
    }
    return out;
}

bool LocalTime(std::tm& t) {
    const std::time_t now = std::time(nullptr);
    return localtime_r(&now, &t) != nullptr;
}

std::string Format(const std::tm& t, const char* fmt) {
    char buf[64];
    return std::strftime(buf, sizeof(buf), fmt, &t) ? std::string(buf) : std::string();
}

std::string MetadataJson(const ne_winner_metadata& m, const uint8_t* letters, const std::string& timestamp) {
    char fitness[32];
    std::snprintf(fitness, sizeof(fitness), "%.9g", double(m.best_fitness));
    std::string j = "{\n";
    j += "  \"format_version\": " + std::to_string(kFormatVersion) + ",\n";
    j += "  \"genome_layout\": \"" + std::string(kGenomeLayout) + "\",\n";
    j += "  \"timestamp_local\": \"" + Escape(timestamp) + "\",\n";
    j += "  \"generation_index\": " + std::to_string(m.generation_index) + ",\n";
    j += "  \"best_fitness\": " + std::string(fitness) + ",\n";
    j += "  \"best_index\": " + std::to_string(m.best_index) + ",\n";
    j += "  \"best_slot_index\": " + std::to_string(m.best_slot_index) + ",\n";
    j += "  \"action_count\": " + std::to_string(m.action_count) + ",\n";
    j += "  \"genome_byte_count\": " + std::to_string(m.genome_byte_count) + ",\n";
    j += "  \"seed\": " + std::to_string(m.seed) + ",\n";
    j += "  \"action_space_path\": \"" + Escape(m.action_space_path) + "\",\n";
    j += "  \"action_space_words\": [";
    for (uint64_t a = 0; a < m.action_count; ++a) {
        if (a) j += ", ";
        j += '"';
        for (int p = 0; p < 5; ++p) j += char('A' + letters[5 * a + p]);
        j += '"';
    }
    j += "]\n}\n";
    return j;
}

bool ValidInputs(const void* genome, const uint8_t* letters, uint64_t words, const ne_winner_metadata* m) {
    if (!genome || !letters || !m || m->action_count == 0 || m->action_count > words ||
        m->action_count > kCatalogCapacity || m->genome_byte_count != ne::SlotStrideBytes(m->action_count) ||
        std::memchr(m->action_space_path, 0, sizeof(m->action_space_path)) == nullptr)
        return false;
    for (uint64_t i = 0; i < 5 * m->action_count; ++i)
        if (letters[i] >= 26) return false;
    return true;
}

bool WriteFile(const std::filesystem::path& path, const void* data, size_t bytes) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data, 1, bytes, f) == bytes;
    return std::fclose(f) == 0 && ok;
}

bool Replace(const std::filesystem::path& tmp, const std::filesystem::path& target) {
    std::error_code ec;
    std::filesystem::rename(tmp, target, ec);
    if (!ec) return true;
    ec.clear();
    std::filesystem::remove(target, ec);
    ec.clear();
    std::filesystem::rename(tmp, target, ec);
    return !ec;
}

// The value after "key": in the metadata, skipping blanks (LocateJsonValue, winner_artifact_reader.cpp:58-76).
bool Locate(const std::string& json, std::string_view key, size_t& pos) {
    const std::string needle = "\"" + std::string(key) + "\"";
    const size_t k = json.find(needle);
    if (k == std::string::npos) return false;
    const size_t colon = json.find(':', k + needle.size());
    if (colon == std::string::npos) return false;
    pos = colon + 1;
    while (pos < json.size() && (json[pos] == ' ' || json[pos] == '\n' || json[pos] == '\r' || json[pos] == '\t'))
        ++pos;
    return pos < json.size();
}

bool ReadString(const std::string& json, std::string_view key, std::string& out) {
    size_t pos = 0;
    if (!Locate(json, key, pos) || json[pos] != '"') return false;
    out.clear();
    for (++pos; pos < json.size(); ++pos) {
        char c = json[pos];
        if (c == '"') return true;
        if (c == '\\') {
            if (++pos >= json.size()) return false;
            switch (json[pos]) {
                case '\\': case '"': c = json[pos]; break;
                case 'n': c = '\n'; break;
                case 'r': c = '\r'; break;
                case 't': c = '\t'; break;
                case 'u': {  // \u00XX: the single-byte code points the writer emits for control bytes
                    if (pos + 4 >= json.size() || json[pos + 1] != '0' || json[pos + 2] != '0') return false;
                    int v = 0;
                    for (int k = 3; k <= 4; ++k) {
                        const char h = json[pos + k];
                        const int d = (h >= '0' && h <= '9') ? h - '0'
                                      : (h >= 'a' && h <= 'f') ? h - 'a' + 10
                                      : (h >= 'A' && h <= 'F') ? h - 'A' + 10 : -1;
                        if (d < 0) return false;
                        v = 16 * v + d;
                    }
                    c = char(v);
                    pos += 4;
                    break;
                }
                default: return false;
            }
        }
        out += c;
    }
    return false;
}

template <typename T>
bool ReadNumber(const std::string& json, std::string_view key, T& out) {
    size_t pos = 0;
    if (!Locate(json, key, pos)) return false;
    const char* begin = json.data() + pos;
    const char* end = begin;
    const char* limit = json.data() + json.size();
    while (end < limit && (*end == '-' || *end == '+' || *end == '.' || (*end >= '0' && *end <= '9') || *end == 'e' ||
                           *end == 'E'))
        ++end;
    if (begin == end) return false;
    if constexpr (std::is_floating_point_v<T>) {
        // from_chars for floats needs a newer libstdc++ than some hosts carry; strtof reads the same text
        const std::string text(begin, end);
        char* stop = nullptr;
        errno = 0;
        const float v = std::strtof(text.c_str(), &stop);
        if (stop != text.c_str() + text.size() || errno == ERANGE) return false;
        out = v;
        return true;
    } else {
        return std::from_chars(begin, end, out).ec == std::errc{};
    }
}

bool ReadWords(const std::string& json, std::string_view key, std::vector<uint8_t>& letters) {
    size_t pos = 0;
    if (!Locate(json, key, pos) || json[pos] != '[') return false;
    letters.clear();
    for (++pos; pos < json.size();) {
        while (pos < json.size() && (json[pos] == ' ' || json[pos] == '\n' || json[pos] == '\r' ||
                                     json[pos] == '\t' || json[pos] == ','))
            ++pos;
        if (pos >= json.size()) return false;
        if (json[pos] == ']') return !letters.empty();
        if (json[pos] != '"' || letters.size() / 5 >= kCatalogCapacity) return false;
        const size_t close = json.find('"', pos + 1);
        if (close == std::string::npos || close - pos - 1 != 5) return false;
        for (size_t i = pos + 1; i < close; ++i) {
            if (json[i] < 'A' || json[i] > 'Z') return false;
            letters.push_back(uint8_t(json[i] - 'A'));
        }
        pos = close + 1;
    }
    return false;
}

}  // namespace

extern "C" {

int ne_write_winner_artifact(const char* binary_path, const char* metadata_path, const void* genome_bytes,
                             const uint8_t* action_letters, uint64_t action_word_count, ne_winner_metadata* metadata) {
    namespace fs = std::filesystem;
    if (!binary_path || !metadata_path || !*binary_path || !*metadata_path ||
        !ValidInputs(genome_bytes, action_letters, action_word_count, metadata))
        return NE_INVALID_ARGUMENT;
    const fs::path bin(binary_path), meta(metadata_path);
    if (bin == meta) return NE_INVALID_ARGUMENT;
    std::error_code ec;
    for (const fs::path& p : {bin, meta})
        if (!p.parent_path().empty()) {
            fs::create_directories(p.parent_path(), ec);
            if (ec) return NE_IO_FAILURE;
        }
    std::tm t{};
    if (!LocalTime(t)) return NE_IO_FAILURE;
    const std::string stamp = Format(t, "%Y-%m-%dT%H:%M:%S");
    const std::string json = MetadataJson(*metadata, action_letters, stamp);
    const fs::path bin_tmp(bin.string() + ".tmp"), meta_tmp(meta.string() + ".tmp");
    const bool ok = WriteFile(bin_tmp, genome_bytes, metadata->genome_byte_count) &&
                    WriteFile(meta_tmp, json.data(), json.size()) && Replace(bin_tmp, bin) && Replace(meta_tmp, meta);
    if (!ok) {
        fs::remove(bin_tmp, ec);
        fs::remove(meta_tmp, ec);
        return NE_IO_FAILURE;
    }
    std::snprintf(metadata->timestamp_local, sizeof(metadata->timestamp_local), "%s", stamp.c_str());
    return NE_OK;
}

int ne_write_winner_artifact_in_directory(const char* directory, const void* genome_bytes,
                                          const uint8_t* action_letters, uint64_t action_word_count,
                                          ne_winner_metadata* metadata, char* binary_path_out,
                                          char* metadata_path_out, uint64_t path_capacity) {
    namespace fs = std::filesystem;
    if (!directory || !binary_path_out || !metadata_path_out || path_capacity == 0 ||
        !ValidInputs(genome_bytes, action_letters, action_word_count, metadata))
        return NE_INVALID_ARGUMENT;
    std::error_code ec;
    fs::create_directories(directory, ec);
    if (ec) return NE_IO_FAILURE;
    std::tm t{};
    if (!LocalTime(t)) return NE_IO_FAILURE;
    // winner-<timestamp>-g<generation>-seed<seed>, then -2, -3, ... (winner_artifact.cpp:79-111, 196-203)
    const std::string stem = "winner-" + Format(t, "%Y-%m-%d_%H-%M-%S") + "-g" +
                             std::to_string(metadata->generation_index) + "-seed" + std::to_string(metadata->seed);
    for (int suffix = 0; suffix < 1000; ++suffix) {
        const std::string name = suffix == 0 ? stem : stem + "-" + std::to_string(suffix + 1);
        const fs::path bin = fs::path(directory) / (name + ".bin"), meta = fs::path(directory) / (name + ".json");
        if (fs::exists(bin) || fs::exists(meta)) continue;
        if (bin.string().size() >= path_capacity || meta.string().size() >= path_capacity) return NE_INVALID_ARGUMENT;
        const int st = ne_write_winner_artifact(bin.c_str(), meta.c_str(), genome_bytes, action_letters,
                                                action_word_count, metadata);
        if (st != NE_OK) return st;
        std::snprintf(binary_path_out, path_capacity, "%s", bin.c_str());
        std::snprintf(metadata_path_out, path_capacity, "%s", meta.c_str());
        return NE_OK;
    }
    return NE_IO_FAILURE;
}

int ne_read_winner_artifact(const char* binary_path, const char* metadata_path, ne_winner_metadata* metadata,
                            void* genome_dst, uint64_t genome_capacity, uint8_t* letters_dst,
                            uint64_t letters_capacity_words) {
    if (!binary_path || !metadata_path || !metadata) return NE_INVALID_ARGUMENT;
    std::ifstream mf(metadata_path, std::ios::binary);
    if (!mf) return NE_IO_FAILURE;
    const std::string json((std::istreambuf_iterator<char>(mf)), std::istreambuf_iterator<char>());
    ne_winner_metadata m{};
    int version = 0;
    std::string layout, stamp, path;
    std::vector<uint8_t> letters;
    if (!ReadNumber(json, "format_version", version) || version != kFormatVersion ||
        !ReadString(json, "genome_layout", layout) || layout != kGenomeLayout ||
        !ReadString(json, "timestamp_local", stamp) || !ReadNumber(json, "generation_index", m.generation_index) ||
        !ReadNumber(json, "best_fitness", m.best_fitness) || !ReadNumber(json, "best_index", m.best_index) ||
        !ReadNumber(json, "best_slot_index", m.best_slot_index) || !ReadNumber(json, "action_count", m.action_count) ||
        !ReadNumber(json, "genome_byte_count", m.genome_byte_count) || !ReadNumber(json, "seed", m.seed) ||
        !ReadString(json, "action_space_path", path) || !ReadWords(json, "action_space_words", letters) ||
        path.size() >= sizeof(m.action_space_path) || stamp.size() >= sizeof(m.timestamp_local))
        return NE_CHECKPOINT_INVALID;
    // IsValidLoadedWinnerArtifact (winner_artifact_reader.cpp:195-201), and this build's slot stride
    std::error_code ec;
    const uint64_t size = std::filesystem::file_size(binary_path, ec);
    if (ec) return NE_IO_FAILURE;
    if (m.action_count == 0 || letters.size() != 5 * m.action_count || m.genome_byte_count == 0 ||
        size != m.genome_byte_count || m.genome_byte_count != ne::SlotStrideBytes(m.action_count))
        return NE_CHECKPOINT_INVALID;
    std::snprintf(m.action_space_path, sizeof(m.action_space_path), "%s", path.c_str());
    std::snprintf(m.timestamp_local, sizeof(m.timestamp_local), "%s", stamp.c_str());
    if (genome_dst) {
        if (genome_capacity < m.genome_byte_count) return NE_INVALID_ARGUMENT;
        std::FILE* f = std::fopen(binary_path, "rb");
        if (!f) return NE_IO_FAILURE;
        const bool ok = std::fread(genome_dst, 1, m.genome_byte_count, f) == m.genome_byte_count;
        std::fclose(f);
        if (!ok) return NE_IO_FAILURE;
    }
    if (letters_dst) {
        if (letters_capacity_words < m.action_count) return NE_INVALID_ARGUMENT;
        std::memcpy(letters_dst, letters.data(), letters.size());
    }
    *metadata = m;
    return NE_OK;
}

}  // extern "C"
