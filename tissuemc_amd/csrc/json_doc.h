// A small JSON document model for the result files: values are built as a tree (objects keep
// their members in insertion order, which is the schema's key order) and then serialized in one
// pass. Numbers are written in the shortest form that reads back as the same double
// (std::to_chars); JSON has no NaN or infinity, so non-finite numbers are written as null.
#pragma once

#include <charconv>
#include <cmath>
#include <cstdint>
#include <memory>
#include <ostream>
#include <string>
#include <utility>
#include <vector>

namespace tmc::json {

class Value {
  public:
    enum class Kind { Null, Bool, Unsigned, Real, Text, List, Record };

    Value() = default;
    Value(bool b) : kind_(Kind::Bool), u_(b ? 1 : 0) {}
    Value(double x) : kind_(Kind::Real), x_(x) {}
    Value(uint64_t u) : kind_(Kind::Unsigned), u_(u) {}
    Value(int i) : kind_(Kind::Unsigned), u_(static_cast<uint64_t>(i < 0 ? 0 : i)) {}
    Value(std::string s) : kind_(Kind::Text), text_(std::move(s)) {}
    Value(const char *s) : kind_(Kind::Text), text_(s) {}

    static Value list() { return Value(Kind::List); }
    static Value record() { return Value(Kind::Record); }
    template <class T> static Value list_of(const std::vector<T> &xs) {
        Value v = list();
        v.items_.reserve(xs.size());
        for (const T &x : xs) v.items_.emplace_back(x);
        return v;
    }

    // Append a member to a record (keys are not checked for uniqueness).
    Value &set(std::string key, Value v) {
        keys_.push_back(std::move(key));
        items_.push_back(std::move(v));
        return *this;
    }
    // Append an element to a list.
    Value &push(Value v) {
        items_.push_back(std::move(v));
        return *this;
    }

    void write(std::ostream &out) const {
        switch (kind_) {
        case Kind::Null: out << "null"; break;
        case Kind::Bool: out << (u_ ? "true" : "false"); break;
        case Kind::Unsigned: out << u_; break;
        case Kind::Real: write_real(out, x_); break;
        case Kind::Text: write_text(out, text_); break;
        case Kind::List:
            out << '[';
            for (size_t i = 0; i < items_.size(); ++i) {
                if (i) out << ',';
                items_[i].write(out);
            }
            out << ']';
            break;
        case Kind::Record:
            out << '{';
            for (size_t i = 0; i < items_.size(); ++i) {
                if (i) out << ',';
                write_text(out, keys_[i]);
                out << ':';
                items_[i].write(out);
            }
            out << '}';
            break;
        }
    }

  private:
    explicit Value(Kind k) : kind_(k) {}

    static void write_real(std::ostream &out, double x) {
        if (!std::isfinite(x)) {
            out << "null";
            return;
        }
        char buf[32];
        const auto r = std::to_chars(buf, buf + sizeof buf, x);
        out.write(buf, r.ptr - buf);
    }

    static void write_text(std::ostream &out, const std::string &s) {
        static const char hex[] = "0123456789abcdef";
        out << '"';
        for (const char ch : s) {
            const auto c = static_cast<unsigned char>(ch);
            if (c == '"' || c == '\\') {
                out << '\\' << ch;
            } else if (c < 0x20) {
                out << "\\u00" << hex[c >> 4] << hex[c & 15];
            } else {
                out << ch;
            }
        }
        out << '"';
    }

    Kind kind_ = Kind::Null;
    uint64_t u_ = 0;
    double x_ = 0.0;
    std::string text_;
    std::vector<std::string> keys_;
    std::vector<Value> items_;
};

}  // namespace tmc::json
