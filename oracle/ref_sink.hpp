// ref_sink.hpp - std::cout redirection for the reference harnesses (TEST
// INFRASTRUCTURE ONLY).
//
// The reference prints from inside OpenMP loops (IPR.cpp:93-102,
// StereoMatch.cpp:102, Shake.cpp:73, STB.cpp:175 ...).  A std::stringbuf put
// under std::cout is then written by several threads at once: its put
// pointer and its reallocation race, writes land past the buffer and the
// heap is corrupted (16 threads writing 7-character numbers through a
// std::ostringstream abort with "stack smashing detected" within seconds).
// LockedSink has no put area, so every character reaches overflow/xsputn,
// which append (or drop) under a mutex.
#pragma once

#include <iostream>
#include <mutex>
#include <streambuf>
#include <string>

class LockedSink : public std::streambuf {
 public:
  explicit LockedSink(bool keep) : keep_(keep) {}
  std::string text() {
    std::lock_guard<std::mutex> lk(mu_);
    return s_;
  }

 protected:
  int_type overflow(int_type c) override {
    if (!traits_type::eq_int_type(c, traits_type::eof()) && keep_) {
      std::lock_guard<std::mutex> lk(mu_);
      s_.push_back(traits_type::to_char_type(c));
    }
    return traits_type::not_eof(c);
  }
  std::streamsize xsputn(const char *p, std::streamsize n) override {
    if (keep_) {
      std::lock_guard<std::mutex> lk(mu_);
      s_.append(p, (size_t)n);
    }
    return n;
  }

 private:
  std::mutex mu_;
  std::string s_;
  bool keep_;
};

// RAII: std::cout into a LockedSink (kept or dropped) for the scope.
struct CoutRedirect {
  LockedSink sink;
  std::streambuf *old;
  explicit CoutRedirect(bool keep) : sink(keep), old(std::cout.rdbuf(&sink)) {}
  ~CoutRedirect() { std::cout.rdbuf(old); }
  std::string text() { return sink.text(); }
};
