//! The reference's builder tests that pin the interface (panic texts, export / import of the
//! affinities, embedding length), on a GPU through libbhtsne_hip.  Run with
//! `BHTSNE_HIP_LIB_DIR=<dir of libbhtsne_hip.so> cargo test` on a machine with an MI355X.
use bhtsne_hip::{Neighbor, TSne};

/// Deterministic samples, as the reference's tests draw them (src/test.rs:973-982).
fn lcg_samples(n: usize, dim: usize, mut state: u64) -> Vec<f32> {
    let mut out = Vec::with_capacity(n * dim);
    for _ in 0..n * dim {
        state = state.wrapping_mul(6364136223846793005).wrapping_add(1442695040888963407);
        out.push(((state >> 33) as f32 / u32::MAX as f32) - 0.5);
    }
    out
}

fn neighbours(n: usize, k: usize) -> Vec<Vec<Neighbor<f32>>> {
    (0..n)
        .map(|i| (1..=k).map(|d| Neighbor { index: (i + d) % n, distance: d as f32 }).collect())
        .collect()
}

#[test]
#[should_panic(expected = "initial embedding has")]
fn warm_start_rejects_wrong_length_barnes_hut() {
    let x = lcg_samples(60, 4, 7);
    let mut t: TSne<2> = TSne::new(&x, 4);
    t.perplexity(5.0).epochs(10).initial_embedding(vec![0.0f32; 7]);
    t.barnes_hut_euclidean(0.5);
}

#[test]
#[should_panic(expected = "same length")]
fn fit_sne_with_neighbors_rejects_ragged_rows() {
    let x = lcg_samples(80, 4, 9);
    let mut nb = neighbours(80, 6);
    nb[3].pop();
    let mut t: TSne<2> = TSne::new(&x, 4);
    t.epochs(10).fit_sne_with_neighbors(&nb);
}

#[test]
#[should_panic(expected = "out of range")]
fn barnes_hut_with_neighbors_rejects_out_of_range_index() {
    let x = lcg_samples(80, 4, 11);
    let mut nb = neighbours(80, 6);
    nb[5][2].index = 80;
    let mut t: TSne<2> = TSne::new(&x, 4);
    t.epochs(10).barnes_hut_with_neighbors(0.5, &nb);
}

#[test]
fn affinities_round_trip_and_embedding_length() {
    let x = lcg_samples(200, 8, 13);
    let mut a: TSne<2> = TSne::new(&x, 8);
    a.perplexity(10.0).epochs(50).fit_sne_euclidean();
    assert_eq!(a.embedding().len(), 400);
    assert!(a.kl_divergence().unwrap().is_finite());
    let aff = a.affinities().unwrap();
    assert_eq!(aff.rows().len(), 201);
    let mut b: TSne<2> = TSne::new(&x, 8);
    b.perplexity(10.0).epochs(50).with_affinities(aff.clone());
    b.fit_sne_euclidean();
    let back = b.affinities().unwrap();
    assert_eq!(back.rows(), aff.rows());
    assert_eq!(back.columns(), aff.columns());
    for (u, v) in back.values().iter().zip(aff.values()) {
        assert!((u - v).abs() <= 1e-6 * v.abs().max(1e-30));
    }
}
