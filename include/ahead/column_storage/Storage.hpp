/*
 * column_storage/Storage.hpp -- device-resident BATs.
 *
 * The reference's BAT<Head,Tail> / TempBAT<Head,Tail> (include/column_storage/Bat.h:54-120,
 * TempBat.h:51-564) hold each side in a ColumnDescriptor (ColumnDescriptor.hpp:43-102): a
 * shared container plus ColumnMetaData; a v2_void_t side has no container and stands for the
 * dense sequence seqbase, seqbase+1, ...  This layer keeps the same shape with the containers
 * in HBM: a `device_vector<T>` owns a hipMalloc'd array, and keeps a host copy only when host
 * code appends to it or iterates over it (each direction is synchronised lazily, on first use
 * from the other side).  mirror_head / reverse / clear_head share containers exactly as in the
 * reference (TempBat.h:157-178), so an operator chain never copies columns between calls.
 */
#ifndef AHEAD_COLUMN_STORAGE_STORAGE_HPP
#define AHEAD_COLUMN_STORAGE_STORAGE_HPP

#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include <column_storage/ColumnMetaData.hpp>

namespace ahead {

    inline void hip_check(hipError_t e, const char *what) {
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
            throw std::bad_alloc();
        if (e != hipSuccess)
            throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
    }

    /* An array in device memory with an optional host copy. */
    template<typename T>
    class device_vector {
        T *d_ = nullptr;
        size_t n_ = 0;
        size_t cap_ = 0;
        mutable std::vector<T> h_;
        mutable bool host_valid_ = true;   // h_ holds the current contents
        mutable bool dev_valid_ = true;    // d_ holds the current contents

        void grow_device(size_t cap) const {
            if (cap <= cap_)
                return;
            T *nd = nullptr;
            hip_check(hipMalloc(reinterpret_cast<void **>(&nd), cap * sizeof(T)), "hipMalloc");
            if (d_ && n_ && dev_valid_)
                hip_check(hipMemcpy(nd, d_, n_ * sizeof(T), hipMemcpyDeviceToDevice), "hipMemcpy");
            if (d_)
                (void) hipFree(d_);
            const_cast<device_vector *>(this)->d_ = nd;
            const_cast<device_vector *>(this)->cap_ = cap;
        }

    public:
        typedef T value_type;

        device_vector() = default;
        explicit device_vector(size_t n) { resize(n); }
        device_vector(const device_vector &) = delete;
        device_vector & operator=(const device_vector &) = delete;
        ~device_vector() {
            if (d_)
                (void) hipFree(d_);
        }

        size_t size() const { return n_; }
        bool empty() const { return n_ == 0; }
        size_t capacity() const { return cap_; }

        /* Device pointer, valid for n_ elements once host-side appends are uploaded. */
        T *data() {
            sync_to_device();
            host_valid_ = false;    // the caller may write through it
            return d_;
        }
        const T *data() const {
            sync_to_device();
            return d_;
        }
        /* Device pointer for a write of exactly `n` elements (no upload of old contents). */
        T *overwrite(size_t n) {
            grow_device(n);
            n_ = n;
            dev_valid_ = true;
            host_valid_ = false;
            h_.clear();
            return d_;
        }
        void reserve(size_t n) {
            if (host_valid_ && !dev_valid_)
                h_.reserve(n);
            else
                grow_device(n);
        }
        void resize(size_t n) {
            sync_to_device();
            grow_device(n);
            n_ = n;
            host_valid_ = false;
        }
        /* Host-side append (loading, tests); the next device use uploads the contents. */
        void push_back(const T &v) {
            sync_to_host();
            h_.push_back(v);
            ++n_;
            dev_valid_ = false;
        }
        void clear() {
            n_ = 0;
            h_.clear();
            host_valid_ = dev_valid_ = true;
        }
        void assign(const T *host, size_t n) {
            overwrite(n);
            if (n)
                hip_check(hipMemcpy(d_, host, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
        }
        const std::vector<T> & host() const {
            sync_to_host();
            return h_;
        }
        void sync_to_device() const {
            if (dev_valid_)
                return;
            grow_device(n_ ? n_ : 1);
            if (n_)
                hip_check(hipMemcpy(d_, h_.data(), n_ * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
            dev_valid_ = true;
        }
        void sync_to_host() const {
            if (host_valid_)
                return;
            h_.resize(n_);
            if (n_)
                hip_check(hipMemcpy(h_.data(), d_, n_ * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
            host_valid_ = true;
        }
    };

    template<typename V2Type>
    struct ColumnDescriptorContainerType {
        typedef device_vector<typename V2Type::type_t> container_t;
    };
    template<>
    struct ColumnDescriptorContainerType<v2_void_t> {
        typedef void container_t;
    };

    template<typename V2Type, typename Container = typename ColumnDescriptorContainerType<V2Type>::container_t>
    class ColumnDescriptor {
    public:
        typedef V2Type v2_type_t;
        typedef Container container_t;
        typedef typename V2Type::type_t type_t;

        std::shared_ptr<container_t> container;
        ColumnMetaData metaData;

        ColumnDescriptor() : container(std::make_shared<container_t>()), metaData(default_metadata<V2Type>()) {}
        explicit ColumnDescriptor(const ColumnMetaData & md) : container(std::make_shared<container_t>()), metaData(md) {}
        ColumnDescriptor(std::shared_ptr<container_t> c, const ColumnMetaData & md) : container(std::move(c)), metaData(md) {}
        ColumnDescriptor(const ColumnDescriptor &) = default;
        ColumnDescriptor & operator=(const ColumnDescriptor &) = default;
    };

    template<>
    class ColumnDescriptor<v2_void_t, void> {
    public:
        typedef v2_void_t v2_type_t;
        typedef void container_t;
        typedef oid_t type_t;

        ColumnMetaData metaData;

        ColumnDescriptor() : metaData(oid_t(0)) {}
        explicit ColumnDescriptor(const ColumnMetaData & md) : metaData(md) {}
        explicit ColumnDescriptor(oid_t seqbase) : metaData(seqbase) {}
    };

    /* Host-side iteration (TempBatIterator.h:49-466): hasNext / head / tail / ++ / position.
     * The first iteration copies the BAT's containers to the host. */
    template<typename Head, typename Tail>
    class BATIterator {
        typedef typename Head::type_t head_t;
        typedef typename Tail::type_t tail_t;
        const std::vector<head_t> *h_ = nullptr;
        const std::vector<tail_t> *t_ = nullptr;
        uint64_t h_seq_ = 0, t_seq_ = 0;
        size_t pos_ = 0, n_ = 0;

    public:
        BATIterator(const std::vector<head_t> *h, uint64_t hs, const std::vector<tail_t> *t, uint64_t ts, size_t n)
                : h_(h), t_(t), h_seq_(hs), t_seq_(ts), n_(n) {}
        bool hasNext() const { return pos_ < n_; }
        BATIterator & operator++() { ++pos_; return *this; }
        void position(size_t p) { pos_ = p; }
        size_t position() const { return pos_; }
        size_t size() const { return n_; }
        head_t head() const { return h_ ? (*h_)[pos_] : static_cast<head_t>(h_seq_ + pos_); }
        tail_t tail() const { return t_ ? (*t_)[pos_] : static_cast<tail_t>(t_seq_ + pos_); }
    };

    template<typename Head, typename Tail>
    class BAT {
    public:
        typedef Head v2_head_t;
        typedef Tail v2_tail_t;
        typedef typename Head::type_t head_t;
        typedef typename Tail::type_t tail_t;
        typedef ColumnDescriptor<Head> coldesc_head_t;
        typedef ColumnDescriptor<Tail> coldesc_tail_t;

        coldesc_head_t head;
        coldesc_tail_t tail;

        BAT() = default;
        BAT(const coldesc_head_t & h, const coldesc_tail_t & t) : head(h), tail(t) {}
        virtual ~BAT() = default;

        /* Number of rows: the head's, or the tail's for a <void, T> BAT (TempBat.h:180-182,
         * :549-551).  A <void, void> BAT has no rows of its own. */
        virtual oid_t size() {
            if constexpr (!std::is_same_v<Head, v2_void_t>)
                return static_cast<oid_t>(head.container->size());
            else if constexpr (!std::is_same_v<Tail, v2_void_t>)
                return static_cast<oid_t>(tail.container->size());
            else
                return 0;
        }
        virtual size_t consumption() {
            size_t c = 0;
            if constexpr (!std::is_same_v<Head, v2_void_t>)
                c += head.container->size() * sizeof(head_t);
            if constexpr (!std::is_same_v<Tail, v2_void_t>)
                c += tail.container->size() * sizeof(tail_t);
            return c;
        }

        virtual BATIterator<Head, Tail> * begin() {
            const std::vector<head_t> *h = nullptr;
            const std::vector<tail_t> *t = nullptr;
            if constexpr (!std::is_same_v<Head, v2_void_t>)
                h = &head.container->host();
            if constexpr (!std::is_same_v<Tail, v2_void_t>)
                t = &tail.container->host();
            return new BATIterator<Head, Tail>(h, head.metaData.seqbase, t, tail.metaData.seqbase, size());
        }

        void reserve(size_t n) {
            if constexpr (!std::is_same_v<Head, v2_void_t>)
                head.container->reserve(n);
            if constexpr (!std::is_same_v<Tail, v2_void_t>)
                tail.container->reserve(n);
        }
        void append(const std::pair<head_t, tail_t> & p) {
            if constexpr (!std::is_same_v<Head, v2_void_t>)
                head.container->push_back(p.first);
            if constexpr (!std::is_same_v<Tail, v2_void_t>)
                tail.container->push_back(p.second);
        }
        void append(tail_t t) {
            static_assert(std::is_same_v<Head, v2_void_t>, "append(tail) is for <void, T> BATs");
            tail.container->push_back(t);
        }

        BAT<Tail, Head> * reverse() { return new BAT<Tail, Head>(coldesc_tail_as_head(), coldesc_head_as_tail()); }
        BAT<Head, Head> * mirror_head() { return new BAT<Head, Head>(head, head); }
        BAT<v2_void_t, Tail> * clear_head() {
            return new BAT<v2_void_t, Tail>(ColumnDescriptor<v2_void_t>(head.metaData), tail);
        }

    private:
        ColumnDescriptor<Tail> coldesc_tail_as_head() { return tail; }
        ColumnDescriptor<Head> coldesc_head_as_tail() { return head; }
    };

    /* The reference distinguishes persistent ColumnBATs from intermediate TempBATs; here both
     * are device-resident arrays, so TempBAT only adds the reference's name and constructors. */
    template<typename Head, typename Tail>
    class TempBAT : public BAT<Head, Tail> {
    public:
        typedef typename BAT<Head, Tail>::coldesc_head_t coldesc_head_t;
        typedef typename BAT<Head, Tail>::coldesc_tail_t coldesc_tail_t;
        TempBAT() = default;
        TempBAT(const coldesc_head_t & h, const coldesc_tail_t & t) : BAT<Head, Tail>(h, t) {}
    };

    /* Builds a <void, V2> BAT from host values (already encoded for v2_res*_t types). */
    template<typename V2>
    TempBAT<v2_void_t, V2> * make_bat(const std::vector<typename V2::type_t> & values, const ColumnMetaData & md = default_metadata<V2>()) {
        auto *b = new TempBAT<v2_void_t, V2>(ColumnDescriptor<v2_void_t>(oid_t(0)), ColumnDescriptor<V2>(md));
        b->tail.container->assign(values.data(), values.size());
        return b;
    }

}

#endif
